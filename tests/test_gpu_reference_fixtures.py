"""GPU parity on the reference's own fixtures: its meshes (python/data/*.obj), its camera
(python/data/cameras.calibration) and its demo SH preset."""
import os

import pytest

from gvv_amd import io, scenes
from parity_util import run_parity

pytestmark = pytest.mark.gpu
DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_data")


def test_cone_demo_render(device):
    """python/test_render.py: cone.obj, batch 2, vertexColor + shaded, 1024^2 (here 512^2)."""
    m = io.read_obj(os.path.join(DATA, "cone.obj"))
    c = io.read_calibration(os.path.join(DATA, "cameras.calibration"), 512, 512)
    s = scenes.from_fixture(m, c, 512, 512, batch=2)
    run_parity(s, device, "vertexColor", "shaded", m2d=True)


@pytest.mark.parametrize("albedo", ["vertexColor", "textured"])
def test_magdalena(device, albedo):
    """python/test_gradients_*.py mesh: magdalena.obj (5118 v / 10115 f)."""
    m = io.read_obj(os.path.join(DATA, "magdalena.obj"), load_texture=False)
    c = io.read_calibration(os.path.join(DATA, "cameras.calibration"), 384, 384)
    s = scenes.from_fixture(m, c, 384, 384, batch=1)
    run_parity(s, device, albedo, "shaded", m2d=albedo == "vertexColor")


def test_monkey_quads(device):
    m = io.read_obj(os.path.join(DATA, "monkey.obj"))
    c = io.read_calibration(os.path.join(DATA, "cameras.calibration"), 256, 256)
    s = scenes.from_fixture(m, c, 256, 256)
    run_parity(s, device, "vertexColor", "shaded", m2d=False)
