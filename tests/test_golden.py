"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py): the oracle must
reproduce them exactly, and the HIP renderer must match them to the parity bar."""
import glob
import os

import numpy as np
import pytest

from conftest import rel_err
from gvv_amd import scenes
from oracle import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(glob.glob(os.path.join(HERE, "*.npz")))


def load(path):
    z = np.load(path, allow_pickle=False)
    s = scenes.Scene(faces=z["faces"], uv=z["uv"], vpos=z["vpos"], vcol=z["vcol"], tex=z["tex"], sh=z["sh"],
                     extr=z["extr"], intr=z["intr"], target=z["target"], width=int(z["width"]), height=int(z["height"]),
                     albedo_mode=str(z["albedo_mode"]), shading_mode=str(z["shading_mode"]))
    return s, z


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_oracle_reproduces_golden(path):
    s, z = load(path)
    orc = O.Oracle(**s.renderer_kwargs())
    out = orc.forward(*s.inputs())
    assert np.array_equal(out["face"], z["out_face"])
    assert np.array_equal(out["bary"].view(np.uint32), z["out_bary"].view(np.uint32))
    assert np.array_equal(out["vnormal"].view(np.uint32), z["out_vnormal"].view(np.uint32))
    assert rel_err(out["render"], z["out_render"]) <= 1e-12
    gr = orc.backward(z["g_render"], z["g_target"], s.target, s.vpos, s.vcol, s.tex, s.sh, out["vnormal"], out["bary"],
                      out["face"], s.extr, s.intr)
    for k in ("pos", "col", "tex", "sh"):
        assert rel_err(gr[k], z["grad_" + k]) <= 1e-12, k


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_gpu_matches_golden(path, device):
    import torch
    from gvv_amd import Rasterizer
    from parity_util import GRAD_TOL, RENDER_TOL, to_dev
    s, z = load(path)
    r = Rasterizer(**s.renderer_kwargs())
    ins = [to_dev(a, device) for a in s.inputs()]
    bary, face, render, vnormal, _ = r.forward(*ins)
    m2d = bool(z["m2d"])
    g = r.backward(to_dev(z["g_render"], device), to_dev(z["g_target"], device) if m2d else None,
                   to_dev(s.target, device), ins[0], ins[1], ins[2], ins[3], vnormal, face, ins[4], ins[5])
    torch.cuda.synchronize()
    assert np.array_equal(face.cpu().numpy(), z["out_face"])
    assert np.array_equal(bary.cpu().numpy().view(np.uint32), z["out_bary"].view(np.uint32))
    assert rel_err(render.cpu().numpy(), z["out_render"]) <= RENDER_TOL
    for k, x in zip(("pos", "col", "tex", "sh"), g):
        e = rel_err(x.cpu().numpy(), z["grad_" + k])
        print(f"{os.path.basename(path)} grad {k} rel err {e:.3e}", flush=True)
        assert e <= GRAD_TOL, k
