"""The C ABI without a GPU: libgvv_amd.so loads, exports every function include/gvv_renderer.h
declares, and rejects invalid arguments on the host before any device call."""
import ctypes
import os
import re

import numpy as np
import pytest

from gvv_amd import _lib
from gvv_amd.renderer import Rasterizer

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gvv_renderer.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?(?:int|void|char|int64_t)\s*\*?\s*(gvv_\w+)\s*\(", text, re.M)))


def test_header_declares_the_abi():
    names = declared_functions()
    for n in ("gvv_create", "gvv_destroy", "gvv_forward", "gvv_backward", "gvv_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    for n in declared_functions():
        assert hasattr(L, n), n
    assert set(_lib.EXPORTS) == set(declared_functions())
    assert L.gvv_version() == 1


def _create(faces, texcoords, n_vertices=3, n_cameras=1, w=8, h=8, albedo=0, shading=0, ifs=1, tfs=1):
    L = _lib.load()
    f = np.ascontiguousarray(np.asarray(faces, np.int32))
    t = np.ascontiguousarray(np.asarray(texcoords, np.float32))
    h_ = ctypes.c_void_p()
    rc = L.gvv_create(f.ctypes.data_as(ctypes.c_void_p), f.size, t.ctypes.data_as(ctypes.c_void_p), t.size,
                      n_vertices, n_cameras, w, h, albedo, shading, ifs, tfs, 0, ctypes.byref(h_))
    return rc, L.gvv_last_error().decode(), h_


@pytest.mark.parametrize("kwargs,msg", [
    (dict(faces=[0, 1], texcoords=[0.0] * 6), "3*F"),
    (dict(faces=[0, 1, 2], texcoords=[0.0] * 5), "6*F"),
    (dict(faces=[0, 1, 3], texcoords=[0.0] * 6), "out of range"),
    (dict(faces=[0, 1, 2], texcoords=[0.0] * 6, n_cameras=0), "number_of_cameras"),
    (dict(faces=[0, 1, 2], texcoords=[0.0] * 6, w=0), "resolution"),
    (dict(faces=[0, 1, 2], texcoords=[0.0] * 6, albedo=7), "albedo"),
    (dict(faces=[0, 1, 2], texcoords=[0.0] * 6, shading=2), "shading"),
    (dict(faces=[0, 1, 2], texcoords=[0.0] * 6, ifs=0), "image_filter_size"),
])
def test_create_rejects_invalid_arguments(kwargs, msg):
    rc, err, h = _create(**kwargs)
    assert rc != 0 and h.value is None
    assert msg in err


def test_null_handle_calls_fail_cleanly():
    L = _lib.load()
    assert L.gvv_forward(None, 1, 1, 1, *([None] * 12)) != 0
    assert "null handle" in L.gvv_last_error().decode()
    assert L.gvv_backward(None, 1, 1, 1, *([None] * 17)) != 0


@pytest.mark.parametrize("albedo,shading", [("bogus", "shaded"), ("textured", "flat")])
def test_python_rejects_mode_strings(albedo, shading):
    with pytest.raises(ValueError):
        Rasterizer([0, 1, 2], [0.0] * 6, 3, 1, 8, 8, albedo_mode=albedo, shading_mode=shading)
