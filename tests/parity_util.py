"""Shared GPU-vs-oracle comparison used by the GPU test modules."""
import numpy as np
import torch

from conftest import elem_rel_stats, rel_err
from gvv_amd import Rasterizer
from oracle import oracle as O

RENDER_TOL = 1e-5
GRAD_TOL = 1e-4


def to_dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def run_parity(scene, device, albedo, shading, m2d, seed=0, image_fs=1, use_bary=False, check_grads=True, tuning=None):
    kw = scene.renderer_kwargs()
    kw.update(albedo_mode=albedo, shading_mode=shading)
    orc = O.Oracle(**kw, image_filter_size=image_fs)
    ref = orc.forward(*scene.inputs())
    rast = Rasterizer(**kw, image_filter_size=image_fs)
    if tuning is not None:
        rast.set_tuning(*tuning)
        assert rast.tuning() == tuple(tuning)
    ins = [to_dev(a, device) for a in scene.inputs()]
    bary, face, render, vnormal, _ = rast.forward(*ins)
    torch.cuda.synchronize()
    face_np = face.cpu().numpy()
    print(f"[{scene.meta.get('name')} {albedo}/{shading}] covered px {int((face_np >= 0).sum())}", flush=True)
    assert np.array_equal(face_np, ref["face"]), f"face ids differ at {int((face_np != ref['face']).sum())} px"
    assert np.array_equal(bary.cpu().numpy().view(np.uint32), ref["bary"].view(np.uint32)), "barycentrics not bit-exact"
    assert np.array_equal(vnormal.cpu().numpy().view(np.uint32), ref["vnormal"].view(np.uint32)), "vertex normals not bit-exact"
    e_render = rel_err(render.cpu().numpy(), ref["render"])
    print(f"    render rel err {e_render:.3e}", flush=True)
    assert e_render <= RENDER_TOL
    if not check_grads:
        return
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    g_render = rng.normal(size=ref["render"].shape).astype(np.float32)
    g_target = rng.normal(size=ref["render"].shape).astype(np.float32) if m2d else None
    gref = orc.backward(g_render, g_target if m2d else np.zeros_like(g_render), scene.target, scene.vpos, scene.vcol,
                        scene.tex, scene.sh, ref["vnormal"], ref["bary"], ref["face"], scene.extr, scene.intr)
    got = rast.backward(to_dev(g_render, device), to_dev(g_target, device) if m2d else None, to_dev(scene.target, device),
                        ins[0], ins[1], ins[2], ins[3], vnormal, face, ins[4], ins[5], bary=bary if use_bary else None)
    torch.cuda.synchronize()
    for name, g in zip(("pos", "col", "tex", "sh"), got):
        g = g.cpu().numpy()
        e = rel_err(g, gref[name])
        emax, emed = elem_rel_stats(g, gref[name])
        print(f"    grad {name:3s} rel err {e:.3e}  (elementwise max {emax:.2e} median {emed:.2e}; |ref|max "
              f"{np.abs(gref[name]).max():.3e})", flush=True)
        assert e <= GRAD_TOL, f"grad {name}: {e}"


