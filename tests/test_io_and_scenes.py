"""CPU tests of the fixture readers (against values in the fixture files' own text) and of the
synthetic workload generators."""
import os

import numpy as np
import pytest

from gvv_amd import io, scenes

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_data")


def test_read_cone_obj():
    m = io.read_obj(os.path.join(DATA, "cone.obj"))
    assert (m.N, m.F) == (5, 6)
    assert np.allclose(m.vertices[1], [2.406190, 1570.561523, -2.628878])
    assert np.allclose(m.colors[1], [0.0, 1.0, 0.0])
    assert m.faces.tolist() == [[0, 1, 2], [2, 1, 4], [4, 1, 3], [3, 1, 0], [4, 0, 2], [4, 3, 0]]
    # per-face expansion of the vt list (OBJReader.py:94-100): "f 1/1/1 2/2/2 3/3/3" and
    # "f 5/6/4 1/7/1 3/8/3" take vt 1,2,3 and vt 6,7,8
    vt = {1: (0.473697, 0.178467), 2: (0.030500, 0.964844), 3: (0.960308, 0.962094),
          6: (0.873816, 0.853557), 7: (0.896875, 0.824535), 8: (0.912500, 0.851995)}
    assert np.allclose(m.texcoords[0], [vt[1], vt[2], vt[3]], atol=1e-6)
    assert np.allclose(m.texcoords[4], [vt[6], vt[7], vt[8]], atol=1e-6)
    assert m.texture is not None and m.texture.shape == (256, 256, 3)
    assert 0.0 <= m.texture.min() and m.texture.max() <= 1.0


def test_read_quads_keep_first_three_corners():
    m = io.read_obj(os.path.join(DATA, "monkey.obj"))
    assert (m.N, m.F) == (504, 499)
    assert m.faces[0].tolist() == [2, 44, 46]          # "f 3/1/3 45/1/45 47/1/47 1/2/1"
    assert m.texture is None                            # its MTL names no image


def test_read_magdalena_topology():
    m = io.read_obj(os.path.join(DATA, "magdalena.obj"), load_texture=False)
    assert (m.N, m.F) == (5118, 10115)
    assert np.unique(m.faces).size == m.N               # no isolated vertex


@pytest.mark.parametrize("size", [(1024, 1024), (512, 256)])
def test_read_calibration_rescale(size):
    c = io.read_calibration(os.path.join(DATA, "cameras.calibration"), *size)
    assert c.C == 1
    su, sv = size[0] / 1024.0, size[1] / 1024.0
    K = c.intrinsics[0].reshape(3, 3)
    assert np.allclose(K, [[752.6881 * su, 0, 523.5678 * su], [0, 752.8311 * sv, 502.7037 * sv], [0, 0, 1]], rtol=1e-6)
    assert np.allclose(c.extrinsics[0][:4], [-0.00292611, 0.0130055, -0.9999111, -734.4974])
    assert np.allclose(c.extrinsics[0][8:], [-0.9970352, -0.07692309, 0.001917182, 3337.621])


def test_workload_generators():
    s0 = scenes.config0()
    assert (s0.N, s0.F, s0.C, s0.width) == (2562, 5120, 4, 256)
    s1 = scenes.config1(with_target=False)
    assert (s1.N, s1.F, s1.C, s1.width, s1.tex.shape[1]) == (50002, 100000, 8, 1024, 1024)
    for s in (s0, s1):
        assert np.unique(s.faces).size == s.N
        assert s.faces.min() >= 0 and s.faces.max() < s.N
    assert 0.0 <= s1.uv.min() and s1.uv.max() <= 1.0
    assert 0.0 <= s1.tex.min() and s1.tex.max() <= 1.0
    # millimetre scale: radius 1 m sphere, cameras 3 m / 3.5 m away
    assert abs(np.linalg.norm(s0.vpos[0], axis=1).mean() - 1000.0) < 5.0
    R = s1.extr[0, 0].reshape(3, 4)[:, :3]
    t = s1.extr[0, 0].reshape(3, 4)[:, 3]
    assert abs(np.linalg.norm(-R.T @ t) - 3500.0) < 1e-2


def test_look_at_convention():
    E = scenes.look_at((0.0, 0.0, -2.0))
    R, t = E[:, :3], E[:, 3]
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-6)
    assert np.allclose(R @ np.array([0.0, 0.0, 0.0]) + t, [0, 0, 2000.0])      # origin 2 m ahead (+z)
    assert (R @ np.array([0.0, 1000.0, 0.0]))[1] < 0                           # world up is image up (-y)


def test_config2_batch_shape():
    batch = scenes.config2(seed=0)
    assert len(batch) == 32
    for s in batch:
        assert 2000 * 0.8 <= s.F <= 200000 * 1.25, s.meta
        assert 1 <= s.C <= 8 and s.width == s.height and s.width in (256, 512, 1024)
        assert s.vpos.shape == (1, s.N, 3) and s.sh.shape == (1, s.C, 27)
        assert s.faces.min() >= 0 and s.faces.max() < s.N
    Fs = sorted(s.F for s in batch)
    assert Fs[-1] / Fs[0] > 10          # genuinely uneven


def test_autotune_cache_file_round_trip(tmp_path):
    """The multi-mesh autotuner's JSON cache: written atomically, read back, and ignored when
    its version is unknown."""
    from gvv_amd import multi
    path = str(tmp_path / "sub" / "autotune.json")
    assert multi._read_cache(path) == {}
    devices = {"gpu|gfx950|256": {"vertexColor/shaded": {"3": {"tile": 8, "large_box": 64}}}}
    multi._write_cache(path, devices)
    assert multi._read_cache(path) == devices
    import json
    with open(path, "w") as fh:
        json.dump({"version": 99, "devices": devices}, fh)
    assert multi._read_cache(path) == {}
