"""Multi-mesh runtime (include/gvv_multi.h): many meshes of uneven size rendered forward and
backward with one hipGraph launch, launch settings chosen per size class by an autotuner.

The reference takes one topology per op instance and loops over the batch on the host
(cpp/src/TensorflowOperators/CudaRenderer/CudaRenderer.cpp:23, :309-328); this has no
counterpart there.
"""
from __future__ import annotations

import ctypes
import json
import os
import tempfile

import numpy as np
import torch

from . import _lib
from .renderer import ALBEDO_MODES, SHADING_MODES

_P = ctypes.c_void_p


class MeshDesc(ctypes.Structure):
    _fields_ = [("faces", _P), ("n_face_idx", ctypes.c_int64), ("texcoords", _P), ("n_vertices", ctypes.c_int32),
                ("n_cameras", ctypes.c_int32), ("width", ctypes.c_int32), ("height", ctypes.c_int32)]


class MeshIO(ctypes.Structure):
    _fields_ = [("batch", ctypes.c_int32), ("tex_h", ctypes.c_int32), ("tex_w", ctypes.c_int32),
                ("vpos", _P), ("vcol", _P), ("tex", _P), ("sh", _P), ("extr", _P), ("intr", _P),
                ("bary", _P), ("face", _P), ("render", _P), ("vnormal", _P),
                ("g_render", _P), ("g_pos", _P), ("g_col", _P), ("g_tex", _P), ("g_sh", _P)]


def _p(t):
    return t.data_ptr() if t is not None else None


def device_key(device) -> str:
    """Autotuner cache key of a GPU: product name, architecture and compute-unit count."""
    props = torch.cuda.get_device_properties(device)
    return f"{props.name}|{getattr(props, 'gcnArchName', '')}|{props.multi_processor_count}"


def _read_cache(path: str) -> dict:
    """The autotuner cache file: {device key: {"albedo/shading": {size class: {"tile", "large_box"}}}}."""
    try:
        with open(path) as fh:
            data = json.load(fh)
    except FileNotFoundError:
        return {}
    if not isinstance(data, dict) or data.get("version") != 1:
        return {}
    return data.get("devices", {})


def _write_cache(path: str, devices: dict):
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=d, prefix=".autotune", suffix=".json")
    with os.fdopen(fd, "w") as fh:
        json.dump({"version": 1, "devices": devices}, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


class MultiRenderer:
    """Owns the runtime handle and, per mesh, device copies of the scene inputs plus output and
    gradient buffers.  `scenes` are gvv_amd.scenes.Scene objects (one mesh each, any batch)."""

    def __init__(self, scenes, albedo_mode="vertexColor", shading_mode="shaded", device=None):
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.scenes = list(scenes)
        self.albedo_mode, self.shading_mode = albedo_mode, shading_mode
        L = _lib.load()
        self._keep = []
        descs = (MeshDesc * len(self.scenes))()
        for d, s in zip(descs, self.scenes):
            f = np.ascontiguousarray(s.faces, np.int32).reshape(-1)
            uv = np.ascontiguousarray(s.uv, np.float32).reshape(-1)
            self._keep += [f, uv]
            d.faces, d.n_face_idx, d.texcoords = f.ctypes.data, f.size, uv.ctypes.data
            d.n_vertices, d.n_cameras, d.width, d.height = s.N, s.C, s.width, s.height
        h = _P()
        _lib.check(L.gvv_multi_create(descs, len(self.scenes), ALBEDO_MODES[albedo_mode], SHADING_MODES[shading_mode],
                                      ctypes.byref(h)), "gvv_multi_create")
        self._h = h
        self._destroy = L.gvv_multi_destroy
        self.io = (MeshIO * len(self.scenes))()
        self.bufs = []
        dev = self.device
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        for io, s in zip(self.io, self.scenes):
            B, C, N, H, W = s.B, s.C, s.N, s.height, s.width
            b = dict(vpos=t(s.vpos), vcol=t(s.vcol), tex=t(s.tex), sh=t(s.sh), extr=t(s.extr.reshape(B, -1)),
                     intr=t(s.intr.reshape(B, -1)),
                     bary=torch.empty((B, C, H, W, 2), device=dev), face=torch.empty((B, C, H, W), dtype=torch.int32, device=dev),
                     render=torch.empty((B, C, H, W, 3), device=dev), vnormal=torch.empty((B, C, N, 3), device=dev),
                     g_render=torch.zeros((B, C, H, W, 3), device=dev), g_pos=torch.empty((B, N, 3), device=dev),
                     g_col=torch.empty((B, N, 3), device=dev), g_tex=torch.empty(s.tex.shape, device=dev),
                     g_sh=torch.empty((B, C, 27), device=dev))
            io.batch, io.tex_h, io.tex_w = B, s.tex.shape[1], s.tex.shape[2]
            for k, v in b.items():
                setattr(io, k, _p(v))
            self.bufs.append(b)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and getattr(self, "_destroy", None) is not None:
            try:
                torch.cuda.synchronize()
            except Exception:      # interpreter shutdown: torch may already be torn down
                pass
            self._destroy(h)
            self._h = None

    @staticmethod
    def _stream():
        return _P(torch.cuda.current_stream().cuda_stream)

    def use_graph(self, on: bool):
        _lib.check(_lib.load().gvv_multi_use_graph(self._h, int(on)), "gvv_multi_use_graph")

    def set_batched(self, on: bool):
        """True: one launch per stage across all meshes; False (default): per-mesh launches on
        parallel graph branches."""
        _lib.check(_lib.load().gvv_multi_set_batched(self._h, int(on)), "gvv_multi_set_batched")

    def run(self, backward: bool = True):
        """Forward of every mesh (+ backward against each mesh's bufs['g_render'])."""
        _lib.check(_lib.load().gvv_multi_run(self._h, self.io, int(backward), self._stream()), "gvv_multi_run")

    def autotune(self, reps: int = 3, cache: str | None = None) -> str:
        """Chooses the launch settings per size class (gvv_multi_autotune).  With `cache`, the path
        of a JSON file of earlier choices keyed by device and mode: when every size class present
        has an entry there it is applied without timing anything ("cache"); otherwise the
        autotuner times the classes ("timed") and the file is updated with its choices."""
        n = len(self.scenes)
        classes = [self.tuning(i)[2] for i in range(n)]
        dev_key = device_key(self.device)
        mode_key = f"{self.albedo_mode}/{self.shading_mode}"
        data = _read_cache(cache) if cache else {}
        entries = data.get(dev_key, {}).get(mode_key, {})
        if cache and all(str(c) in entries for c in set(classes)):
            for i, c in enumerate(classes):
                e = entries[str(c)]
                self.set_tuning(i, e["tile"], e["large_box"])
            return "cache"
        _lib.check(_lib.load().gvv_multi_autotune(self._h, self.io, int(reps), self._stream()), "gvv_multi_autotune")
        if cache:
            for i, c in enumerate(classes):
                tile, box, _ = self.tuning(i)
                entries[str(c)] = {"tile": tile, "large_box": box}
            data.setdefault(dev_key, {})[mode_key] = entries
            _write_cache(cache, data)
        return "timed"

    def set_tuning(self, i: int, tile: int, large_box: int):
        """Backward tile edge and large-face threshold of mesh i."""
        _lib.check(_lib.load().gvv_multi_set_tuning(self._h, int(i), int(tile), int(large_box)), "gvv_multi_set_tuning")

    def tuning(self, i: int):
        """(tile, large_box, size_class) of mesh i."""
        t, lb, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _lib.check(_lib.load().gvv_multi_tuning(self._h, i, ctypes.byref(t), ctypes.byref(lb), ctypes.byref(c)),
                   "gvv_multi_tuning")
        return t.value, lb.value, c.value

    def pixels(self):
        return sum(s.B * s.C * s.width * s.height for s in self.scenes)

    def triangles(self):
        return sum(s.B * s.C * s.F for s in self.scenes)
