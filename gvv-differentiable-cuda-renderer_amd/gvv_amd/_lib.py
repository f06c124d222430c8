"""ctypes binding of libgvv_amd.so (C ABI declared in include/gvv_renderer.h,
include/gvv_multi.h, include/gvv_proposal.h and include/gvv_fit.h).

The library is the only compute path: if it has not been built this module raises; there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("GVV_AMD_LIB", os.path.join(_PKG, "libgvv_amd.so"))

EXPORTS = ("gvv_last_error", "gvv_version", "gvv_create", "gvv_destroy", "gvv_forward", "gvv_backward",
           "gvv_backward_noalloc",
           "gvv_profile_enable", "gvv_profile_report", "gvv_set_tuning", "gvv_get_tuning",
           "gvv_set_depth_pass", "gvv_get_depth_pass",
           "gvv_calibrate_atomics",
           "gvv_conv_out_size", "gvv_conv_forward", "gvv_conv_backward_data", "gvv_conv_workspace_floats",
           "gvv_conv_backward_weight", "gvv_conv_prepare_weights", "gvv_pack_mosaic",
           "gvv_multi_create", "gvv_multi_destroy", "gvv_multi_run", "gvv_multi_use_graph", "gvv_multi_autotune",
           "gvv_multi_set_batched", "gvv_multi_set_tuning",
           "gvv_multi_tuning",
           "gvv_masked_sq_workspace_floats", "gvv_masked_sq_error", "gvv_masked_sq_error_grad",
           "gvv_shape_forward", "gvv_shape_workspace_floats", "gvv_shape_backward")

_lib = None


class GvvError(RuntimeError):
    pass


def load():
    """Load and type the HIP renderer library; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GvvError(f"{LIB_PATH} is missing: build it with `python gvv-differentiable-cuda-renderer_amd/build.py` "
                       "(or __graft_entry__.build()); the renderer has no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    P, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    L.gvv_last_error.restype = ctypes.c_char_p
    L.gvv_version.restype = ctypes.c_int
    L.gvv_create.argtypes = [P, i64, P, i64, i32, i32, i32, i32, i32, i32, i32, i32, i32, ctypes.POINTER(P)]
    L.gvv_create.restype = ctypes.c_int
    L.gvv_destroy.argtypes = [P]
    L.gvv_destroy.restype = None
    L.gvv_forward.argtypes = [P, i32, i32, i32] + [P] * 12
    L.gvv_forward.restype = ctypes.c_int
    L.gvv_backward.argtypes = [P, i32, i32, i32] + [P] * 17
    L.gvv_backward.restype = ctypes.c_int
    L.gvv_backward_noalloc.argtypes = [P, i32, i32, i32] + [P] * 17
    L.gvv_backward_noalloc.restype = ctypes.c_int
    L.gvv_profile_enable.argtypes = [P, i32]
    L.gvv_profile_enable.restype = ctypes.c_int
    L.gvv_profile_report.argtypes = [P, ctypes.c_char_p, i64]
    L.gvv_profile_report.restype = i64
    L.gvv_calibrate_atomics.argtypes = [P, i64, i64, i32, P]
    L.gvv_calibrate_atomics.restype = ctypes.c_int
    L.gvv_set_tuning.argtypes = [P, i32, i32]
    L.gvv_set_tuning.restype = ctypes.c_int
    L.gvv_get_tuning.argtypes = [P, ctypes.POINTER(i32), ctypes.POINTER(i32)]
    L.gvv_get_tuning.restype = ctypes.c_int
    L.gvv_set_depth_pass.argtypes = [P, i32]
    L.gvv_set_depth_pass.restype = ctypes.c_int
    L.gvv_get_depth_pass.argtypes = [P, ctypes.POINTER(i32)]
    L.gvv_get_depth_pass.restype = ctypes.c_int
    L.gvv_conv_out_size.argtypes = [i32] * 4
    L.gvv_conv_out_size.restype = ctypes.c_int
    L.gvv_conv_forward.argtypes = [i32] * 8 + [P, P, P, i32, P, i32, P]
    L.gvv_conv_forward.restype = ctypes.c_int
    L.gvv_conv_backward_data.argtypes = [i32] * 8 + [P, P, P, P, P]
    L.gvv_conv_backward_data.restype = ctypes.c_int
    L.gvv_conv_workspace_floats.argtypes = [i32] * 8
    L.gvv_conv_workspace_floats.restype = i64
    L.gvv_conv_backward_weight.argtypes = [i32] * 8 + [P, P, P, P, P, i64, P]
    L.gvv_conv_backward_weight.restype = ctypes.c_int
    L.gvv_conv_prepare_weights.argtypes = [i32, i32, i32, P, P, P, P]
    L.gvv_conv_prepare_weights.restype = ctypes.c_int
    L.gvv_pack_mosaic.argtypes = [P, i32, i32, i32, i32, i32, i32, i32, P, P]
    L.gvv_pack_mosaic.restype = ctypes.c_int
    L.gvv_multi_create.argtypes = [P, i32, i32, i32, ctypes.POINTER(P)]
    L.gvv_multi_create.restype = ctypes.c_int
    L.gvv_multi_destroy.argtypes = [P]
    L.gvv_multi_destroy.restype = None
    L.gvv_multi_run.argtypes = [P, P, i32, P]
    L.gvv_multi_run.restype = ctypes.c_int
    L.gvv_multi_use_graph.argtypes = [P, i32]
    L.gvv_multi_use_graph.restype = ctypes.c_int
    L.gvv_multi_set_batched.argtypes = [P, i32]
    L.gvv_multi_set_batched.restype = ctypes.c_int
    L.gvv_multi_set_tuning.argtypes = [P, i32, i32, i32]
    L.gvv_multi_set_tuning.restype = ctypes.c_int
    L.gvv_multi_autotune.argtypes = [P, P, i32, P]
    L.gvv_multi_autotune.restype = ctypes.c_int
    L.gvv_multi_tuning.argtypes = [P, i32, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32)]
    L.gvv_multi_tuning.restype = ctypes.c_int
    L.gvv_masked_sq_workspace_floats.argtypes = [i32, i64]
    L.gvv_masked_sq_workspace_floats.restype = i64
    L.gvv_masked_sq_error.argtypes = [i32, i64, P, P, P, P, P, P]
    L.gvv_masked_sq_error.restype = ctypes.c_int
    L.gvv_masked_sq_error_grad.argtypes = [i32, i64, P, P, P, P, P, P]
    L.gvv_masked_sq_error_grad.restype = ctypes.c_int
    f32 = ctypes.c_float
    L.gvv_shape_forward.argtypes = [i32, i32, i32, P, P, f32] + [P] * 7
    L.gvv_shape_forward.restype = ctypes.c_int
    L.gvv_shape_workspace_floats.argtypes = [i32, i32, i32]
    L.gvv_shape_workspace_floats.restype = i64
    L.gvv_shape_backward.argtypes = [i32, i32, i32, P, P, P, P, f32, P, P, P, P]
    L.gvv_shape_backward.restype = ctypes.c_int
    _lib = L
    return L


def check(rc: int, what: str):
    if rc != 0:
        raise GvvError(f"{what}: {load().gvv_last_error().decode()}")
