"""ctypes front-end of the CPU oracle (``oracle/gvv_oracle.cpp``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py``.  The renderer product never imports this module.

The oracle restates the reference rasterizer (cpp/src/Renderer/CUDABasedRasterization.cu,
cpp/src/Renderer/CUDABasedRasterizationGrad.cu, cpp/src/Utils/RendererUtil.h) with a binary32
visibility stage and FP64 shading/gradients; see the header of gvv_oracle.cpp.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgvv_oracle.so")

ALBEDO = {"vertexColor": 0, "textured": 1, "normal": 2, "lighting": 3, "foregroundMask": 4}
SHADING = {"shaded": 0, "shadeless": 1}

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle shared library with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "libgvv_oracle.so"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i32, i64 = ctypes.c_int32, ctypes.c_int64
        L.gvvo_last_error.restype = ctypes.c_char_p
        L.gvvo_set_threads.argtypes = [ctypes.c_int]
        L.gvvo_set_threads.restype = ctypes.c_int
        L.gvvo_create.argtypes = [P, i64, P, i64, i32, i32, i32, i32, i32, i32, i32, i32, i32,
                                  ctypes.POINTER(ctypes.c_void_p)]
        L.gvvo_create.restype = ctypes.c_int
        L.gvvo_destroy.argtypes = [P]
        L.gvvo_forward.argtypes = [P, i32, i32, i32] + [P] * 11
        L.gvvo_forward.restype = ctypes.c_int
        L.gvvo_backward.argtypes = [P, i32, i32, i32] + [P] * 16
        L.gvvo_backward.restype = ctypes.c_int
        L.gvvo_reference_vertex_faces.argtypes = [P, i64, i32, P, i64, P, i64, ctypes.POINTER(ctypes.c_int64)]
        L.gvvo_reference_vertex_faces.restype = ctypes.c_int
        L.gvvo_camera.argtypes = [P] * 5
        L.gvvo_pixel_barycentric.argtypes = [P, P, ctypes.c_float, ctypes.c_float, P, P, P]
        L.gvvo_bary_jacobian.argtypes = [P] * 4
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def set_threads(n: int) -> int:
    return lib().gvvo_set_threads(int(n))


class Oracle:
    """One oracle handle per (topology, attributes), like one reference op instance."""

    def __init__(self, faces, texcoords, n_vertices, n_cameras, width, height,
                 albedo_mode="textured", shading_mode="shaded", image_filter_size=1,
                 texture_filter_size=1, compute_normal_map=False):
        self.faces = np.ascontiguousarray(np.asarray(faces, dtype=np.int32).reshape(-1))
        self.texcoords = _f32(texcoords).reshape(-1)
        self.N, self.C, self.W, self.H = int(n_vertices), int(n_cameras), int(width), int(height)
        self.albedo_mode, self.shading_mode = albedo_mode, shading_mode
        self.compute_normal_map = bool(compute_normal_map)
        h = ctypes.c_void_p()
        rc = lib().gvvo_create(_p(self.faces), self.faces.size, _p(self.texcoords), self.texcoords.size,
                               self.N, self.C, self.W, self.H, ALBEDO[albedo_mode], SHADING[shading_mode],
                               int(image_filter_size), int(texture_filter_size), int(compute_normal_map),
                               ctypes.byref(h))
        if rc != 0:
            raise ValueError(lib().gvvo_last_error().decode())
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            lib().gvvo_destroy(self._h)
            self._h = None

    def forward(self, vpos, vcol, tex, sh, extr, intr):
        vpos, vcol, tex, sh, extr, intr = map(_f32, (vpos, vcol, tex, sh, extr, intr))
        B, texH, texW = tex.shape[0], tex.shape[1], tex.shape[2]
        C, H, W, N = self.C, self.H, self.W, self.N
        bary = np.empty((B, C, H, W, 2), np.float32)
        face = np.empty((B, C, H, W), np.int32)
        render = np.empty((B, C, H, W, 3), np.float64)
        vnormal = np.empty((B, C, N, 3), np.float32)
        nmap = np.zeros((B, texH, texW, 3), np.float64)
        rc = lib().gvvo_forward(self._h, B, texH, texW, _p(vpos), _p(vcol), _p(tex), _p(sh), _p(extr), _p(intr),
                                _p(bary), _p(face), _p(render), _p(vnormal), _p(nmap))
        if rc != 0:
            raise ValueError(lib().gvvo_last_error().decode())
        return dict(bary=bary, face=face, render=render, vnormal=vnormal, normal_map=nmap)

    def backward(self, g_render, g_target, target, vpos, vcol, tex, sh, vnormal, bary, face, extr, intr):
        g_render, g_target, target, vpos, vcol, tex, sh, vnormal, bary, extr, intr = map(
            _f32, (g_render, g_target, target, vpos, vcol, tex, sh, vnormal, bary, extr, intr))
        face = np.ascontiguousarray(np.asarray(face, dtype=np.int32))
        B, texH, texW = vpos.shape[0], tex.shape[1], tex.shape[2]
        g_pos = np.empty((B, self.N, 3), np.float64)
        g_col = np.empty((B, self.N, 3), np.float64)
        g_tex = np.empty((B, texH, texW, 3), np.float64)
        g_sh = np.empty((B, self.C, 27), np.float64)
        rc = lib().gvvo_backward(self._h, B, texH, texW, _p(g_render), _p(g_target), _p(target), _p(vpos),
                                 _p(vcol), _p(tex), _p(sh), _p(vnormal), _p(bary), _p(face), _p(extr), _p(intr),
                                 _p(g_pos), _p(g_col), _p(g_tex), _p(g_sh))
        if rc != 0:
            raise ValueError(lib().gvvo_last_error().decode())
        return dict(pos=g_pos, col=g_col, tex=g_tex, sh=g_sh)


def reference_vertex_faces(faces, n_vertices):
    """Vertex->face lists exactly as CUDABasedRasterization.cpp:125-154 builds them."""
    faces = np.ascontiguousarray(np.asarray(faces, dtype=np.int32).reshape(-1))
    lst = np.empty(faces.size, np.int32)
    pairs = np.empty(2 * int(n_vertices), np.int32)
    n_list = ctypes.c_int64()
    n = lib().gvvo_reference_vertex_faces(_p(faces), faces.size, int(n_vertices), _p(lst), lst.size,
                                          _p(pairs), pairs.size, ctypes.byref(n_list))
    return lst[: n_list.value].copy(), pairs[: 2 * n].reshape(n, 2).copy()


def camera(extr12, intr9):
    e, k = _f32(extr12), _f32(intr9)
    invE, invP, o = np.empty(16, np.float32), np.empty(16, np.float32), np.empty(3, np.float32)
    lib().gvvo_camera(_p(e), _p(k), _p(invE), _p(invP), _p(o))
    return invE.reshape(4, 4), invP.reshape(4, 4), o


def pixel_barycentric(extr12, intr9, px, py, tri9):
    e, k, v = _f32(extr12), _f32(intr9), _f32(tri9)
    d, abc = np.empty(3, np.float32), np.empty(3, np.float32)
    lib().gvvo_pixel_barycentric(_p(e), _p(k), ctypes.c_float(px), ctypes.c_float(py), _p(v), _p(d), _p(abc))
    return d, abc


def bary_jacobian(o3, d3, tri9):
    o, d, v = _f32(o3), _f32(d3), _f32(tri9)
    out = np.empty(27, np.float64)
    lib().gvvo_bary_jacobian(_p(o), _p(d), _p(v), _p(out))
    return out.reshape(3, 9)
