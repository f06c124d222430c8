"""gvv_backward_noalloc (the backward with its scratch in the library's fixed device array)
against gvv_backward on the same handle and inputs, called through the C ABI: two calls in a
row, the second with a smaller batch, in the textured and the vertex-colour modes (with and
without the model-to-data term), plus the size limit and the zero-gradient modes."""
import ctypes

import numpy as np
import pytest
import torch

from conftest import rel_err
from gvv_amd import Rasterizer, _lib, scenes

pytestmark = pytest.mark.gpu


def _noalloc():
    return _lib.load().gvv_backward_noalloc


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _scene(batch):
    s = scenes.random_soup(31, n_faces=400, n_cameras=3, width=96, height=80, batch=batch)
    s.uv = np.random.Generator(np.random.PCG64(5)).uniform(0, 1, size=s.uv.shape).astype(np.float32)
    s.tex = np.ascontiguousarray(np.broadcast_to(scenes.smooth_texture(np.random.Generator(np.random.PCG64(6)), 40, 56),
                                                 (batch, 40, 56, 3)), np.float32)
    return s


def _call_both(rast, s, device, m2d, seed):
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
    vpos, vcol, tex, sh, extr, intr = (t(a) for a in s.inputs())
    bary, face, render, vnormal, _ = rast.forward(vpos, vcol, tex, sh, extr, intr)
    rng = np.random.Generator(np.random.PCG64(seed))
    g_render = t(rng.normal(size=tuple(render.shape)).astype(np.float32))
    g_target = t(rng.normal(size=tuple(render.shape)).astype(np.float32)) if m2d else None
    target = t(s.target)
    ref = rast.backward(g_render, g_target, target, vpos, vcol, tex, sh, vnormal, face, extr, intr)
    B = s.B
    outs = [torch.full(r.shape, float("nan"), device=device) for r in ref]
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = _noalloc()(rast._h, B, tex.shape[1], tex.shape[2], _p(g_render), _p(g_target), _p(target), _p(vpos),
                    _p(vcol), _p(tex), _p(sh), _p(vnormal), None, _p(face), _p(extr.reshape(B, -1)),
                    _p(intr.reshape(B, -1)), *(_p(o) for o in outs), stream)
    _lib.check(rc, "gvv_backward_noalloc")
    torch.cuda.synchronize()
    return ref, outs


@pytest.mark.parametrize("albedo,m2d", [("textured", False), ("vertexColor", True), ("foregroundMask", False)])
def test_noalloc_matches_backward_twice_second_smaller(device, albedo, m2d):
    for batch, seed in ((3, 1), (1, 2)):
        s = _scene(batch)
        kw = s.renderer_kwargs()
        kw.update(albedo_mode=albedo)
        rast = Rasterizer(**kw) if batch == 3 else rast      # one handle for both calls
        ref, outs = _call_both(rast, s, device, m2d, seed)
        for name, r, o in zip(("pos", "col", "tex", "sh"), ref, outs):
            r, o = r.cpu().numpy(), o.cpu().numpy()
            assert np.isfinite(o).all(), name
            e = rel_err(o, r)
            print(f"{albedo} batch {batch} grad {name}: rel diff {e:.2e}", flush=True)
            assert e <= 1e-5, (albedo, batch, name, e)


def test_noalloc_zero_modes_and_size_limit(device):
    s = _scene(1)
    kw = s.renderer_kwargs()
    kw.update(albedo_mode="normal")
    rast = Rasterizer(**kw)
    ref, outs = _call_both(rast, s, device, False, 3)
    for o in outs:
        assert torch.count_nonzero(o) == 0
    # a handle whose scratch would exceed the fixed array is refused with a message
    big = scenes.random_soup(32, n_faces=20, n_cameras=64, width=2048, height=2048)
    r2 = Rasterizer(**big.renderer_kwargs())
    r2.set_tuning(8, 256)
    B = 1
    g = torch.zeros(1, device=device)
    rc = _noalloc()(r2._h, B, 1, 1, *([_p(g)] * 17), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    msg = _lib.load().gvv_last_error().decode()
    print("refused:", msg, flush=True)
    assert rc != 0 and "scratch" in msg
