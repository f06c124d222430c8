"""The C ABI without a GPU: libgvv_amd.so loads, exports every function include/gvv_renderer.h,
include/gvv_multi.h, include/gvv_proposal.h and include/gvv_fit.h declare, and rejects invalid arguments on the host before any
device call."""
import ctypes
import os
import re

import numpy as np
import pytest

from gvv_amd import _lib
from gvv_amd.renderer import Rasterizer

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("gvv_renderer.h", "gvv_multi.h", "gvv_proposal.h", "gvv_fit.h")]


def declared_functions():
    text = "".join(open(h).read() for h in HEADERS)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?(?:int|void|char|int64_t)\s*\*?\s*(gvv_\w+)\s*\(", text, re.M)))


def test_header_declares_the_abi():
    names = declared_functions()
    for n in ("gvv_create", "gvv_destroy", "gvv_forward", "gvv_backward", "gvv_last_error",
              "gvv_conv_forward", "gvv_conv_backward_data", "gvv_conv_backward_weight"):
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


def test_conv_rejects_invalid_geometry():
    L = _lib.load()
    P = ctypes.c_void_p(16)     # never dereferenced: validation fails first
    assert L.gvv_conv_out_size(256, 3, 2, 1) == 128
    assert L.gvv_conv_out_size(4, 1, 1, 0) == 4
    rc = L.gvv_conv_forward(2, 8, 8, 3, 32, 3, 2, 1, P, P, P, 1, P, 0, None)
    assert rc != 0 and "multiples of 8" in L.gvv_last_error().decode()
    rc = L.gvv_conv_forward(2, 2, 2, 8, 32, 5, 1, 0, P, P, P, 1, P, 0, None)
    assert rc != 0 and "kernel larger" in L.gvv_last_error().decode()
    rc = L.gvv_conv_forward(2, 8, 8, 8, 32, 3, 2, 1, ctypes.c_void_p(18), P, P, 1, P, 0, None)
    assert rc != 0 and "aligned" in L.gvv_last_error().decode()
    need = L.gvv_conv_workspace_floats(2, 8, 8, 8, 32, 3, 2, 1)
    assert need > 0
    rc = L.gvv_conv_backward_weight(2, 8, 8, 8, 32, 3, 2, 1, P, P, P, P, P, need - 1, None)
    assert rc != 0 and "workspace" in L.gvv_last_error().decode()
    assert L.gvv_conv_workspace_floats(2, 8, 8, 7, 32, 3, 2, 1) == -1


def test_multi_structs_match_the_c_layout():
    """gvv_mesh_desc / gvv_mesh_io as declared in include/gvv_multi.h (natural alignment)."""
    from gvv_amd.multi import MeshDesc, MeshIO
    assert ctypes.sizeof(MeshDesc) == 8 + 8 + 8 + 4 * 4
    assert ctypes.sizeof(MeshIO) == 3 * 4 + 4 + 15 * 8
    assert MeshIO.vpos.offset == 16 and MeshIO.g_sh.offset == 16 + 14 * 8
    text = open(os.path.join(ROOT, "include", "gvv_multi.h")).read()
    io_fields = re.findall(r"\*\s*(\w+);", text[text.index("typedef struct gvv_mesh_io"):])
    assert io_fields[:15] == [f for f, _ in MeshIO._fields_[3:]]


def test_multi_rejects_modes_without_gradients():
    L = _lib.load()
    h = ctypes.c_void_p()
    from gvv_amd.multi import MeshDesc
    d = (MeshDesc * 1)()
    assert L.gvv_multi_create(d, 1, 2, 0, ctypes.byref(h)) != 0
    assert "vertexColor, textured and foregroundMask" in L.gvv_last_error().decode()
    assert L.gvv_multi_create(d, 0, 0, 0, ctypes.byref(h)) != 0


def test_renderer_cache_key_separates_streams_and_devices():
    """The op-level renderer cache keys on the HIP stream and device as well as the attributes:
    work on two streams gets two handles (a handle's scratch is for one stream at a time)."""
    from gvv_amd.renderer import cache_key
    faces, tc = [0, 1, 2], [0.0] * 6
    args = (faces, tc, 3, 1, 8, 8, "vertexColor", "shaded", 2, 2, False)
    k = cache_key(*args, 0, 1234)
    assert k == cache_key(*args, 0, 1234)
    assert k != cache_key(*args, 0, 5678)
    assert k != cache_key(*args, 1, 1234)
    assert k != cache_key(faces, tc, 3, 1, 8, 8, "textured", "shaded", 2, 2, False, 0, 1234)
    assert k != cache_key([0, 2, 1], tc, 3, 1, 8, 8, "vertexColor", "shaded", 2, 2, False, 0, 1234)


def test_masked_sq_error_rejects_invalid_sizes_on_the_host():
    L = _lib.load()
    assert L.gvv_masked_sq_workspace_floats(0, 10) == -1
    assert L.gvv_masked_sq_workspace_floats(2, 0) == -1
    assert L.gvv_masked_sq_workspace_floats(2, 4096) == 2
    assert L.gvv_masked_sq_workspace_floats(3, 4097) == 6
    assert L.gvv_masked_sq_error(0, 10, None, None, None, None, None, None) != 0
    assert "subjects" in L.gvv_last_error().decode()
    assert L.gvv_masked_sq_error(1, 10, None, None, None, None, None, None) != 0
    assert "null" in L.gvv_last_error().decode()
    assert L.gvv_masked_sq_error_grad(1, 0, None, None, None, None, None, None) != 0


def test_shape_kernels_reject_invalid_sizes_on_the_host():
    L = _lib.load()
    assert L.gvv_shape_workspace_floats(2, 100, 0) == -1
    assert L.gvv_shape_workspace_floats(2, 100, 129) == -1
    assert L.gvv_shape_workspace_floats(2, 100, 8) == (2 + 1) * 2 * 20
    assert L.gvv_shape_forward(1, 10, 200, None, None, 1.0, *([None] * 7)) != 0
    assert "basis" in L.gvv_last_error().decode()
    assert L.gvv_shape_backward(1, 10, 8, None, None, None, None, 1.0, None, None, None, None) != 0
    assert "null" in L.gvv_last_error().decode()
