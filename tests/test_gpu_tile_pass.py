"""The screen-tile depth pass (faces binned per screen tile of 16x16, 32x16 or 32x32 pixels, LDS
z-buffer, resolve fused at tile end) against the oracle on the cases that exercise its edges: every albedo/shading mode, tiles
cut by the image border (tiny and ragged images), faces crossing the camera plane, sub-pixel and
duplicate faces, near-full-screen faces and faces spanning many tiles (the large-face pass and
its tile marks), crowded bins, wide images; and equality with the face-binned pass."""
import numpy as np
import pytest
import torch

from gvv_amd import Rasterizer, scenes
from parity_util import run_parity, to_dev

pytestmark = pytest.mark.gpu
TILED = 1
SHAPES = [1, 2, 3]            # 16x16, 32x16, 32x32 pixel tiles


@pytest.mark.parametrize("albedo", ["vertexColor", "textured", "normal", "lighting", "foregroundMask"])
def test_single_triangle_modes(device, albedo):
    for shading in ("shaded", "shadeless"):
        run_parity(scenes.single_triangle(), device, albedo, shading, m2d=True, depth_pass=TILED)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_stress_visibility(device, seed, shape):
    s = scenes.stress_soup(seed)
    run_parity(s, device, "vertexColor" if seed % 2 == 0 else "textured", "shaded", m2d=seed == 3,
               check_grads=seed < 2, depth_pass=shape)


@pytest.mark.parametrize("seed", [0, 1])
def test_soups(device, seed):
    s = scenes.random_soup(seed, n_faces=120, n_cameras=2, width=64, height=48, batch=2)
    run_parity(s, device, "textured", "shaded", m2d=True, seed=seed, depth_pass=TILED)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("hw", [(1, 1), (3, 5), (17, 9), (33, 47)])
def test_tiny_and_ragged_images(device, hw, shape):
    h, w = hw
    s = scenes.random_soup(13, n_faces=50, n_cameras=2, width=w, height=h)
    run_parity(s, device, "textured", "shaded", m2d=True, depth_pass=shape)


def test_icosphere_config0(device):
    run_parity(scenes.config0(), device, "vertexColor", "shaded", m2d=False, depth_pass=TILED)


@pytest.mark.parametrize("shape", SHAPES)
def test_cylinder_reduced_config1(device, shape):
    s = scenes.config1(width=256, height=256, cols=100, rows=80, n_cameras=2, tex_size=256)
    run_parity(s, device, "textured", "shaded", m2d=False, depth_pass=shape)


@pytest.mark.parametrize("shape", SHAPES)
def test_many_faces_per_tile(device, shape):
    """5,000 faces in a 96x96 view: many bins overflow their capacity, so faces also take the
    large-face pass, and that pass's keys merge with the LDS z-buffers."""
    s = scenes.random_soup(11, n_faces=5000, n_cameras=2, width=96, height=96)
    run_parity(s, device, "vertexColor", "shaded", m2d=False, depth_pass=shape)


def test_wide_image_and_huge_face(device):
    s = scenes.config1(width=8300, height=64, cols=120, rows=40, n_cameras=1, tex_size=64)
    run_parity(s, device, "textured", "shaded", m2d=False, depth_pass=TILED)
    s = scenes.single_triangle(width=8300, height=2100, z_m=0.8,
                               tri_m=((-2.0, -1.0, 0.0), (2.0, -1.0, 0.0), (0.0, 1.2, 0.0)))
    run_parity(s, device, "vertexColor", "shaded", m2d=False, check_grads=False, depth_pass=TILED)


@pytest.mark.parametrize("shape", SHAPES)
def test_equal_to_face_binned_pass_over_repeated_calls(device, shape):
    """Both depth passes give identical outputs, also on repeated calls of one handle (the
    tile counts, marks and global keys are reset by the calls that consume them)."""
    s = scenes.config1(width=320, height=200, cols=120, rows=90, n_cameras=3, tex_size=128)
    ins = [to_dev(a, device) for a in s.inputs()]
    ref = Rasterizer(**s.renderer_kwargs())
    out0 = [t.clone() for t in ref.forward(*ins)[:4]]
    tiled = Rasterizer(**s.renderer_kwargs())
    tiled.set_depth_pass(shape)
    for _ in range(3):
        out1 = tiled.forward(*ins)[:4]
        torch.cuda.synchronize()
        for a, b in zip(out0, out1):
            assert torch.equal(a, b)
