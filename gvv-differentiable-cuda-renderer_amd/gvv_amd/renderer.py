"""Host side of the MI355X rasterizer: handle ownership, the op-level functions and the
``torch.autograd.Function`` that connects them.

Mirrors the reference's TensorFlow op pair and gradient registration
(python/CudaRenderer.py:79-99 and :170-215; cpp/src/TensorflowOperators/CudaRenderer/CudaRenderer.cpp:5-33,
CudaRendererGrad.cpp:6-39).  Tensors are PyTorch-ROCm device tensors; all compute is done by
libgvv_amd.so on the caller's current HIP stream.
"""
from __future__ import annotations

import ctypes
import hashlib
import threading

import numpy as np
import torch

from . import _lib

ALBEDO_MODES = {"vertexColor": 0, "textured": 1, "normal": 2, "lighting": 3, "foregroundMask": 4}
SHADING_MODES = {"shaded": 0, "shadeless": 1}


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream_ptr(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


class Rasterizer:
    """One renderer instance per (topology, attributes), the analogue of one reference op kernel
    instance (CudaRenderer::CudaRenderer, CudaRenderer.cpp:37-155)."""

    def __init__(self, faces, texcoords, n_vertices, n_cameras, width, height, albedo_mode="textured",
                 shading_mode="shaded", image_filter_size=1, texture_filter_size=1, compute_normal_map=False):
        if albedo_mode not in ALBEDO_MODES:
            raise ValueError(f"invalid albedo mode {albedo_mode!r}; expected one of {sorted(ALBEDO_MODES)}")
        if shading_mode not in SHADING_MODES:
            raise ValueError(f"invalid shading mode {shading_mode!r}; expected one of {sorted(SHADING_MODES)}")
        faces = np.ascontiguousarray(np.asarray(faces, dtype=np.int32).reshape(-1))
        texcoords = np.ascontiguousarray(np.asarray(texcoords, dtype=np.float32).reshape(-1))
        self.N, self.C, self.W, self.H = int(n_vertices), int(n_cameras), int(width), int(height)
        self.F = faces.size // 3
        self.albedo_mode, self.shading_mode = albedo_mode, shading_mode
        # foregroundMask forces shadeless (CudaRenderer.cpp:72-76)
        self.effective_shading = "shadeless" if albedo_mode == "foregroundMask" else shading_mode
        self.image_filter_size, self.texture_filter_size = int(image_filter_size), int(texture_filter_size)
        self.compute_normal_map = bool(compute_normal_map)
        self._h = None
        self._zero = {}
        L = _lib.load()
        h = ctypes.c_void_p()
        rc = L.gvv_create(faces.ctypes.data_as(ctypes.c_void_p), faces.size, texcoords.ctypes.data_as(ctypes.c_void_p),
                          texcoords.size, self.N, self.C, self.W, self.H, ALBEDO_MODES[albedo_mode],
                          SHADING_MODES[shading_mode], self.image_filter_size, self.texture_filter_size,
                          int(self.compute_normal_map), ctypes.byref(h))
        if rc != 0:
            raise ValueError(L.gvv_last_error().decode())
        self._h = h
        self._destroy = L.gvv_destroy      # kept: module globals may be gone at interpreter exit
        self.faces_np = faces.reshape(-1, 3)

    def __del__(self):
        h = getattr(self, "_h", None)
        destroy = getattr(self, "_destroy", None)
        if h and destroy is not None:
            try:
                torch.cuda.synchronize()
            except Exception:      # interpreter shutdown: torch may already be torn down
                pass
            destroy(h)
            self._h = None

    # -- shapes ------------------------------------------------------------------------------
    def _check_inputs(self, vpos, vcol, tex, sh, extr, intr, B):
        C, N = self.C, self.N
        expect = {"vertex_pos": (vpos, (B, N, 3)), "vertex_color": (vcol, (B, N, 3)), "sh_coeff": (sh, (B, C, 27)),
                  "extrinsics": (extr, (B, C * 12)), "intrinsics": (intr, (B, C * 9))}
        for name, (t, shape) in expect.items():
            if tuple(t.shape) != shape:
                raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {shape}")
        if tex.dim() != 4 or tex.shape[0] != B or tex.shape[3] != 3:
            raise ValueError(f"texture has shape {tuple(tex.shape)}, expected ({B}, texH, texW, 3)")

    @staticmethod
    def _prep(t, device, dtype=torch.float32):
        t = torch.as_tensor(t, device=device)
        if t.dtype != dtype:
            t = t.to(dtype)
        return t.contiguous()

    # -- forward / backward on raw tensors ----------------------------------------------------
    def forward(self, vpos, vcol, tex, sh, extr, intr):
        """Returns (barycentric_buffer, face_buffer, render_buffer, vertex_normal, normal_map)."""
        dev = vpos.device if torch.is_tensor(vpos) else torch.device("cuda")
        if dev.type != "cuda":
            raise ValueError("the renderer runs on a HIP device; got tensors on " + str(dev))
        vpos, vcol, tex, sh = (self._prep(x, dev) for x in (vpos, vcol, tex, sh))
        B = tex.shape[0]                                    # batch from texture dim 0 (CudaRenderer.cpp:207)
        extr = self._prep(extr, dev).reshape(B, -1)
        intr = self._prep(intr, dev).reshape(B, -1)
        self._check_inputs(vpos, vcol, tex, sh, extr, intr, B)
        texH, texW = tex.shape[1], tex.shape[2]
        C, H, W, N = self.C, self.H, self.W, self.N
        bary = torch.empty((B, C, H, W, 2), device=dev, dtype=torch.float32)
        face = torch.empty((B, C, H, W), device=dev, dtype=torch.int32)
        render = torch.empty((B, C, H, W, 3), device=dev, dtype=torch.float32)
        vnormal = torch.empty((B, C, N, 3), device=dev, dtype=torch.float32)
        if self.compute_normal_map:
            nmap = torch.empty((B, texH, texW, 3), device=dev, dtype=torch.float32)
        else:
            # one zero scalar per device, made on first use: no fill per call (which a stream
            # capture of the call would otherwise record)
            z = self._zero.get(dev)
            if z is None:
                z = self._zero[dev] = torch.zeros((), device=dev, dtype=torch.float32)
            nmap = z.expand(B, texH, texW, 3)
        rc = _lib.load().gvv_forward(self._h, B, texH, texW, _ptr(vpos), _ptr(vcol), _ptr(tex), _ptr(sh), _ptr(extr),
                                     _ptr(intr), _ptr(bary), _ptr(face), _ptr(render), _ptr(vnormal),
                                     _ptr(nmap) if self.compute_normal_map else None, _stream_ptr(dev))
        _lib.check(rc, "gvv_forward")
        return bary, face, render, vnormal, nmap

    def backward(self, g_render, g_target, target, vpos, vcol, tex, sh, vnormal, face, extr, intr, bary=None,
                 vertex_color_grad=True):
        """Returns (vertex_pos_grad, vertex_color_grad, texture_grad, sh_coeff_grad).

        ``g_target`` may be None (zero: the model-to-data term is skipped); ``bary`` may be None
        (barycentrics are recomputed from the face ids, bit-identical to the forward's);
        ``vertex_color_grad=False`` skips the vertex-colour gradient (returned as None)."""
        dev = vpos.device
        vpos, vcol, tex, sh, vnormal = (self._prep(x, dev) for x in (vpos, vcol, tex, sh, vnormal))
        B = vpos.shape[0]                                   # batch from vertex_pos dim 0 (CudaRendererGrad.cpp:195)
        extr = self._prep(extr, dev).reshape(B, -1)
        intr = self._prep(intr, dev).reshape(B, -1)
        face = self._prep(face, dev, torch.int32)
        C, H, W, N = self.C, self.H, self.W, self.N
        if g_render is None:
            g_render = torch.zeros((B, C, H, W, 3), device=dev, dtype=torch.float32)
        g_render = self._prep(g_render, dev)
        if g_target is not None:
            g_target = self._prep(g_target, dev)
            target = self._prep(target, dev)
            if tuple(g_target.shape) != (B, C, H, W, 3) or tuple(target.shape) != (B, C, H, W, 3):
                raise ValueError("target_buffer_grad and target_image must be (B, C, H, W, 3)")
        else:
            target = None
        if bary is not None:
            bary = self._prep(bary, dev)
        if tuple(g_render.shape) != (B, C, H, W, 3) or tuple(face.shape) != (B, C, H, W):
            raise ValueError("render_buffer_grad / face_buffer shapes do not match the renderer")
        if tuple(vnormal.shape) != (B, C, N, 3):
            raise ValueError(f"vertex_normal has shape {tuple(vnormal.shape)}, expected {(B, C, N, 3)}")
        texH, texW = tex.shape[1], tex.shape[2]
        g_pos = torch.empty((B, N, 3), device=dev, dtype=torch.float32)
        g_col = torch.empty((B, N, 3), device=dev, dtype=torch.float32) if vertex_color_grad else None
        g_tex = torch.empty((B, texH, texW, 3), device=dev, dtype=torch.float32)
        g_sh = torch.empty((B, C, 27), device=dev, dtype=torch.float32)
        rc = _lib.load().gvv_backward(self._h, B, texH, texW, _ptr(g_render), _ptr(g_target), _ptr(target), _ptr(vpos),
                                      _ptr(vcol), _ptr(tex), _ptr(sh), _ptr(vnormal), _ptr(bary), _ptr(face),
                                      _ptr(extr), _ptr(intr), _ptr(g_pos), _ptr(g_col), _ptr(g_tex), _ptr(g_sh),
                                      _stream_ptr(dev))
        _lib.check(rc, "gvv_backward")
        return g_pos, g_col, g_tex, g_sh

    def set_tuning(self, tile: int = 16, large_box: int = 256):
        """Backward pixel-tile edge (8/16/32) and the raster's large-face threshold in pixels."""
        _lib.check(_lib.load().gvv_set_tuning(self._h, int(tile), int(large_box)), "gvv_set_tuning")

    def tuning(self):
        t, lb = ctypes.c_int32(), ctypes.c_int32()
        _lib.check(_lib.load().gvv_get_tuning(self._h, ctypes.byref(t), ctypes.byref(lb)), "gvv_get_tuning")
        return t.value, lb.value

    def set_depth_pass(self, mode: int):
        """0: faces binned by bounding-box size with a global visibility buffer; 1: screen tiles
        binned in LDS with the resolve fused at tile end.  Both give identical results."""
        _lib.check(_lib.load().gvv_set_depth_pass(self._h, int(mode)), "gvv_set_depth_pass")

    def depth_pass(self) -> int:
        m = ctypes.c_int32()
        _lib.check(_lib.load().gvv_get_depth_pass(self._h, ctypes.byref(m)), "gvv_get_depth_pass")
        return m.value

    def profile(self, enable: bool = True):
        """Enable/disable per-kernel HIP-event timing of this handle (clears the figures)."""
        _lib.check(_lib.load().gvv_profile_enable(self._h, int(enable)), "gvv_profile_enable")

    def profile_report(self) -> dict:
        """{kernel: (total_ms, launches)} accumulated since profiling was enabled."""
        L = _lib.load()
        n = L.gvv_profile_report(self._h, None, 0)
        buf = ctypes.create_string_buffer(int(n) + 1)
        L.gvv_profile_report(self._h, buf, len(buf))
        out = {}
        for line in buf.value.decode().splitlines():
            name, ms, cnt = line.split()
            out[name] = (float(ms), int(cnt))
        return out

    def render(self, vpos, vcol, tex, sh, target, extr, intr):
        """Differentiable forward: returns the six op outputs as autograd-tracked tensors."""
        return RenderFunction.apply(self, vpos, vcol, tex, sh, target, extr, intr)


class RenderFunction(torch.autograd.Function):
    """Forward + registered gradient of the reference op (python/CudaRenderer.py:170-215).

    Gradients flow to vertex_pos, vertex_color, texture and sh_coeff from render_buffer and
    target_image_out; the incoming gradients of barycentric_buffer, face_buffer,
    vertex_normal and normal_map are ignored, and target_image, extrinsics and intrinsics
    receive zero gradient, as in the reference."""

    @staticmethod
    def forward(ctx, rast, vpos, vcol, tex, sh, target, extr, intr):
        ctx.set_materialize_grads(False)
        bary, face, render, vnormal, nmap = rast.forward(vpos, vcol, tex, sh, extr, intr)
        ctx.rast = rast
        ctx.save_for_backward(vpos, vcol, tex, sh, target, extr, intr, vnormal, face)
        ctx.mark_non_differentiable(face)
        # target_image_out is the target input itself (the reference copies it, :286)
        return bary, face, render, vnormal, target, nmap

    @staticmethod
    def backward(ctx, g_bary, g_face, g_render, g_vnormal, g_target, g_nmap):
        vpos, vcol, tex, sh, target, extr, intr, vnormal, face = ctx.saved_tensors
        rast = ctx.rast
        if g_render is None and g_target is None:
            grads = (None, None, None, None)
        else:
            grads = rast.backward(g_render, g_target, target, vpos, vcol, tex, sh, vnormal, face, extr, intr,
                                  vertex_color_grad=ctx.needs_input_grad[2])
        need = ctx.needs_input_grad
        out = [None]
        for i, g in enumerate(grads):
            out.append(g if need[1 + i] else None)
        for i, t in ((5, target), (6, extr), (7, intr)):
            out.append(torch.zeros_like(t) if need[i] else None)
        return tuple(out)


# ---------------------------------------------------------------------------------------------
# op-level functions (the analogues of customOperators.cuda_renderer_gpu / _grad_gpu)
# ---------------------------------------------------------------------------------------------
_cache: dict = {}
_cache_lock = threading.Lock()
_CACHE_MAX = 32


def _digest(a):
    return hashlib.blake2b(np.ascontiguousarray(a).tobytes(), digest_size=16).hexdigest()


def cache_key(faces, texture_coordinates, number_of_vertices, number_of_cameras, render_resolution_u,
              render_resolution_v, albedo_mode, shading_mode, image_filter_size, texture_filter_size,
              compute_normal_map, device_index, stream_id):
    """Key of the renderer cache: the op attributes (TF caches one op kernel per attr set) plus
    the device and the HIP stream.  A handle's scratch (visibility keys, cameras, gradient
    partials) may only be used by one stream at a time, so work queued on two streams -- which
    the host-side lock of a handle cannot order on the GPU -- gets two handles."""
    faces = np.asarray(faces, dtype=np.int32).reshape(-1)
    tc = np.asarray(texture_coordinates, dtype=np.float32).reshape(-1)
    return (_digest(faces), _digest(tc), int(number_of_vertices), int(number_of_cameras), int(render_resolution_u),
            int(render_resolution_v), albedo_mode, shading_mode, int(image_filter_size), int(texture_filter_size),
            bool(compute_normal_map), int(device_index), int(stream_id))


def get_rasterizer(faces, texture_coordinates, number_of_vertices, number_of_cameras, render_resolution_u,
                   render_resolution_v, albedo_mode, shading_mode, image_filter_size=2, texture_filter_size=2,
                   compute_normal_map=False) -> Rasterizer:
    """Returns the cached renderer for these attributes on the current device and stream."""
    dev = torch.cuda.current_device()
    key = cache_key(faces, texture_coordinates, number_of_vertices, number_of_cameras, render_resolution_u,
                    render_resolution_v, albedo_mode, shading_mode, image_filter_size, texture_filter_size,
                    compute_normal_map, dev, torch.cuda.current_stream(dev).cuda_stream)
    with _cache_lock:
        r = _cache.get(key)
        if r is None:
            r = Rasterizer(faces, texture_coordinates, number_of_vertices, number_of_cameras, render_resolution_u,
                           render_resolution_v, albedo_mode, shading_mode, image_filter_size, texture_filter_size,
                           compute_normal_map)
            if len(_cache) >= _CACHE_MAX:
                _cache.pop(next(iter(_cache)))
            _cache[key] = r
        return r


def cuda_renderer_gpu(vertex_pos, vertex_color, texture, sh_coeff, target_image, extrinsics, intrinsics, *, faces,
                      texture_coordinates, number_of_vertices, number_of_cameras, render_resolution_u=512,
                      render_resolution_v=512, albedo_mode, shading_mode, image_filter_size=2,
                      texture_filter_size=2, compute_normal_map=False, name=None):
    """Forward op with the reference signature (CudaRenderer.cpp:5-33, defaults included).
    Returns (barycentric_buffer, face_buffer, render_buffer, vertex_normal, target_image_out,
    normal_map)."""
    rast = get_rasterizer(faces, texture_coordinates, number_of_vertices, number_of_cameras, render_resolution_u,
                          render_resolution_v, albedo_mode, shading_mode, image_filter_size, texture_filter_size,
                          compute_normal_map)
    dev = vertex_pos.device if torch.is_tensor(vertex_pos) else torch.device("cuda", torch.cuda.current_device())
    target_image = Rasterizer._prep(target_image, dev)
    return rast.render(vertex_pos, vertex_color, texture, sh_coeff, target_image, extrinsics, intrinsics)


def cuda_renderer_grad_gpu(render_buffer_grad, vertex_pos, vertex_color, texture, sh_coeff, target_image,
                           vertex_normal, barycentric_buffer, face_buffer, target_buffer_grad, extrinsics,
                           intrinsics, *, faces, texture_coordinates, number_of_vertices, number_of_cameras,
                           render_resolution_u=512, render_resolution_v=512, albedo_mode, shading_mode,
                           image_filter_size=2, texture_filter_size=2, name=None):
    """Gradient op with the reference signature (CudaRendererGrad.cpp:6-39).  Returns
    (vertex_pos_grad, vertex_color_grad, texture_grad, sh_coeff_grad)."""
    if albedo_mode not in ("vertexColor", "textured", "foregroundMask"):
        raise ValueError(f"the gradient op accepts vertexColor, textured or foregroundMask, not {albedo_mode!r}")
    rast = get_rasterizer(faces, texture_coordinates, number_of_vertices, number_of_cameras, render_resolution_u,
                          render_resolution_v, albedo_mode, shading_mode, image_filter_size, texture_filter_size)
    return rast.backward(render_buffer_grad, target_buffer_grad, target_image, vertex_pos, vertex_color, texture,
                         sh_coeff, vertex_normal, face_buffer, extrinsics, intrinsics, bary=barycentric_buffer)
