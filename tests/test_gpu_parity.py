"""GPU parity: the HIP renderer (through its C ABI) against the CPU oracle on the same seeded
inputs.  Bar: face ids, barycentrics and vertex normals bit-exact; render and normal map within
1e-5 and gradients within 1e-4 relative (max-norm, rel_err in conftest) of the FP64 oracle."""
import numpy as np
import pytest
import torch

from conftest import rel_err
from gvv_amd import Rasterizer, scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu

from parity_util import GRAD_TOL, RENDER_TOL, run_parity as _run, to_dev as _t  # noqa: E402


@pytest.mark.parametrize("albedo", ["vertexColor", "textured", "normal", "lighting", "foregroundMask"])
@pytest.mark.parametrize("shading", ["shaded", "shadeless"])
def test_single_triangle_modes(device, albedo, shading):
    _run(scenes.single_triangle(), device, albedo, shading, m2d=True)


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("albedo", ["vertexColor", "textured", "foregroundMask"])
def test_soup_modes(device, seed, albedo):
    s = scenes.random_soup(seed, n_faces=120, n_cameras=2, width=64, height=48, batch=2)
    _run(s, device, albedo, "shaded", m2d=seed != 1, seed=seed, use_bary=seed == 2)


def test_soup_shadeless_filter2(device):
    s = scenes.random_soup(5, n_faces=80, n_cameras=3, width=40, height=40)
    _run(s, device, "vertexColor", "shadeless", m2d=True, image_fs=2)


def test_icosphere_config0(device):
    _run(scenes.config0(), device, "vertexColor", "shaded", m2d=False)


def test_icosphere_config0_textured_m2d(device):
    s = scenes.config0(width=128, height=128, level=3)
    s.tex = scenes.smooth_texture(np.random.Generator(np.random.PCG64(3)), 64, 48)[None]
    s.uv = np.random.Generator(np.random.PCG64(4)).uniform(0, 1, size=s.uv.shape).astype(np.float32)
    _run(s, device, "textured", "shaded", m2d=True)


def test_cylinder_reduced_config1(device):
    s = scenes.config1(width=256, height=256, cols=100, rows=80, n_cameras=2, tex_size=256)
    _run(s, device, "textured", "shaded", m2d=False)


@pytest.mark.parametrize("tex_hw", [(64, 48), (37, 53)])
def test_normal_map(device, tex_hw):
    s = scenes.random_soup(7, n_faces=60, n_cameras=2, width=32, height=32, tex_size=tex_hw)
    kw = s.renderer_kwargs()
    orc = O.Oracle(**kw, compute_normal_map=True)
    ref = orc.forward(*s.inputs())
    rast = Rasterizer(**kw, compute_normal_map=True)
    bary, face, render, vnormal, nmap = rast.forward(*[_t(a, device) for a in s.inputs()])
    torch.cuda.synchronize()
    # normal-map mode skips rasterisation: the pixel outputs keep their initial values
    assert (face.cpu().numpy() == -1).all()
    assert np.array_equal(render.cpu().numpy()[..., 1], np.ones(render.shape[:-1], np.float32))
    assert np.array_equal(vnormal.cpu().numpy().view(np.uint32), ref["vnormal"].view(np.uint32))
    e = rel_err(nmap.cpu().numpy(), ref["normal_map"])
    print(f"    normal map rel err {e:.3e}", flush=True)
    assert e <= RENDER_TOL


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_stress_visibility(device, seed):
    """Camera-plane crossings, near-full-screen, edge-on, sub-pixel and duplicate faces."""
    s = scenes.stress_soup(seed)
    _run(s, device, "vertexColor" if seed % 2 == 0 else "textured", "shaded", m2d=seed == 3, check_grads=seed < 2)


@pytest.mark.parametrize("with_near", [False, True])
def test_vertex_behind_the_camera(device, with_near):
    """A vertex behind the camera (clamped projection depth) and a triangle 0.01 mm in front of
    the camera: the oracle pin's scene (test_oracle_pins.py), on the GPU.  Without the
    model-to-data term: its projection Jacobian divides by the near triangle's camera depth,
    0.01 mm formed as the difference of 2 m coordinates, which binary32 holds to about 2%."""
    _run(scenes.behind_camera(with_near), device, "textured", "shaded", m2d=False)


@pytest.mark.parametrize("tuning", [(8, 64), (32, 1024), (8, 16), (32, 256)])
def test_tuning_settings(device, tuning):
    """Every backward tile edge and large-face threshold the autotuner may pick keeps the
    forward bit-exact and the gradients within tolerance."""
    s = scenes.config0(width=160, height=120, level=3)
    s.tex = scenes.smooth_texture(np.random.Generator(np.random.PCG64(3)), 64, 48)[None]
    s.uv = np.random.Generator(np.random.PCG64(4)).uniform(0, 1, size=s.uv.shape).astype(np.float32)
    _run(s, device, "vertexColor", "shaded", m2d=True, tuning=tuning)
    _run(s, device, "textured", "shaded", m2d=False, tuning=tuning)


def test_large_face_queue_longer_than_the_persistent_grid(device):
    """More queued large faces than the large-face kernel's persistent wavefronts (4096): with a
    16 px threshold nearly every face of this soup is queued (2 views x 5000 faces), so every
    wavefront takes several in its fixed stride."""
    s = scenes.random_soup(11, n_faces=5000, n_cameras=2, width=96, height=96)
    _run(s, device, "vertexColor", "shaded", m2d=False, tuning=(16, 16))


def test_nothing_visible(device):
    """A mesh far outside every view: no pixel is covered, the buffers hold the background and
    every gradient is zero, as in the oracle."""
    s = scenes.random_soup(12, n_faces=60, n_cameras=2, width=40, height=24)
    s.vpos = s.vpos + np.float32(5.0e6)
    for albedo in ("vertexColor", "textured"):
        _run(s, device, albedo, "shaded", m2d=True)


@pytest.mark.parametrize("hw", [(1, 1), (3, 5), (17, 9)])
def test_tiny_and_ragged_images(device, hw):
    """Images smaller than one backward tile and not a multiple of any tile edge."""
    h, w = hw
    s = scenes.random_soup(13, n_faces=50, n_cameras=2, width=w, height=h)
    for tuning in ((8, 64), (32, 16)):
        _run(s, device, "textured", "shaded", m2d=True, tuning=tuning)


def test_wide_images_unpacked_candidates(device):
    """Images more than 8192 px wide keep the unpacked candidate words (slot | pixel index)."""
    s = scenes.config1(width=8300, height=64, cols=120, rows=40, n_cameras=1, tex_size=64)
    _run(s, device, "textured", "shaded", m2d=False)


def test_face_box_beyond_24_bits(device):
    """One face covering an 8300 x 2100 view: its box holds 17.4 M pixels, past 2^24, so the
    large-face pass needs all 32 bits of the candidate word for the pixel index."""
    s = scenes.single_triangle(width=8300, height=2100, z_m=0.8,
                               tri_m=((-2.0, -1.0, 0.0), (2.0, -1.0, 0.0), (0.0, 1.2, 0.0)))
    _run(s, device, "vertexColor", "shaded", m2d=False, check_grads=False)


@pytest.mark.parametrize("m2d", [False, True])
def test_backward_without_vertex_colour_gradient(device, m2d):
    """g_col = NULL (Rasterizer.backward(vertex_color_grad=False)): the vertexColor backward
    skips the colour terms; position, texture and SH gradients agree with the full call within
    the gradient tolerance (the face slots then hold fewer terms, so float atomics add in
    another order), and the no-allocation entry point accepts it too."""
    s = scenes.config0(width=128, height=128, level=3)
    kw = s.renderer_kwargs()
    rast = Rasterizer(**kw)
    ins = [_t(a, device) for a in s.inputs()]
    vpos, vcol, tex, sh, extr, intr = ins
    bary, face, render, vnormal, _ = rast.forward(*ins)
    g = torch.randn(render.shape, generator=torch.Generator(device="cpu").manual_seed(5)).to(device)
    gt = tgt = None
    if m2d:
        gt = torch.randn(render.shape, generator=torch.Generator(device="cpu").manual_seed(6)).to(device)
        tgt = torch.rand(render.shape, generator=torch.Generator(device="cpu").manual_seed(7)).to(device)
    full = rast.backward(g, gt, tgt, vpos, vcol, tex, sh, vnormal, face, extr, intr)
    part = rast.backward(g, gt, tgt, vpos, vcol, tex, sh, vnormal, face, extr, intr, vertex_color_grad=False)
    assert part[1] is None
    for name, a, b in (("g_pos", full[0], part[0]), ("g_tex", full[2], part[2]), ("g_sh", full[3], part[3])):
        e = rel_err(b.cpu().numpy(), a.cpu().numpy())
        print(f"no colour gradient m2d={m2d}: {name} rel err {e:.2e}", flush=True)
        assert e <= GRAD_TOL
    assert float(full[1].abs().max()) > 0.0
    # the no-allocation entry point with g_col = NULL
    import ctypes
    from gvv_amd import _lib
    L = _lib.load()
    B, N = vpos.shape[0], vpos.shape[1]
    gp = torch.empty((B, N, 3), device=device)
    gx = torch.empty_like(part[2])
    gs = torch.empty_like(part[3])
    p = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
    rc = L.gvv_backward_noalloc(rast._h, B, tex.shape[1], tex.shape[2], p(g), p(gt), p(tgt), p(vpos), p(vcol), p(tex),
                                p(sh), p(vnormal), None, p(face), p(extr.reshape(B, -1)), p(intr.reshape(B, -1)),
                                p(gp), None, p(gx), p(gs), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    _lib.check(rc, "gvv_backward_noalloc")
    torch.cuda.synchronize()
    assert rel_err(gp.cpu().numpy(), full[0].cpu().numpy()) <= GRAD_TOL
