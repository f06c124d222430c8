"""Readers for the reference's on-disk formats.

* OBJ with per-vertex colours on ``v`` lines, triangle faces ``f v/vt/vn``, per-face
  texture-coordinate expansion and an MTL ``map_Kd`` texture (reference
  python/utils/OBJReader.py:41-100,199-214).  The texture is read with PIL instead of cv2;
  channels are RGB in [0, 1] as after the reference's BGR->RGB conversion and /255.
* Skeletool ``.calibration`` camera files (reference python/utils/CameraReader.py:3-49):
  extrinsics are the first 12 values of the 4x4 ``extrinsic`` row; intrinsics keep the 3x3
  part of the 3x4 ``intrinsic`` row (every 4th value skipped) and are rescaled from the
  calibrated ``size`` to the render resolution.

Differences from the reference readers: a ``v`` line without colours gets colour 1 (the
reference raises), and an ``mtllib`` name without a leading "./" is used as written (the
reference strips its first two characters unconditionally).
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np


@dataclass
class ObjMesh:
    vertices: np.ndarray          # (N, 3) float32
    colors: np.ndarray            # (N, 3) float32
    faces: np.ndarray             # (F, 3) int32, 0-based
    texcoords: np.ndarray         # (F, 3, 2) float32, per face corner
    texture: np.ndarray | None    # (H, W, 3) float32 in [0, 1]
    mtl_path: str | None

    @property
    def N(self):
        return self.vertices.shape[0]

    @property
    def F(self):
        return self.faces.shape[0]


def read_obj(path: str, load_texture: bool = True) -> ObjMesh:
    folder = os.path.dirname(path)
    verts, cols, vt, fv, ft = [], [], [], [], []
    mtl = None
    with open(path) as fh:
        for line in fh:
            sp = line.split()
            if not sp:
                continue
            if sp[0] == "v":
                verts.append([float(x) for x in sp[1:4]])
                cols.append([float(x) for x in sp[4:7]] if len(sp) >= 7 else [1.0, 1.0, 1.0])
            elif sp[0] == "vt":
                vt.append([float(sp[1]), float(sp[2])])
            elif sp[0] == "f":
                corners = [c.split("/") for c in sp[1:4]]
                fv.append([int(c[0]) - 1 for c in corners])
                ft.append([int(c[1]) - 1 if len(c) > 1 and c[1] else -1 for c in corners])
            elif sp[0] == "mtllib":
                name = sp[1][2:] if sp[1].startswith("./") else sp[1]
                mtl = os.path.join(folder, name)
    faces = np.asarray(fv, np.int32).reshape(-1, 3)
    vt = np.asarray(vt, np.float32).reshape(-1, 2)
    ft = np.asarray(ft, np.int64).reshape(-1, 3)
    if vt.shape[0] and (ft >= 0).all():
        texcoords = vt[ft]
    else:
        texcoords = np.zeros((faces.shape[0], 3, 2), np.float32)
    texture = None
    if load_texture and mtl and os.path.exists(mtl):
        texture = read_mtl_texture(mtl)
    return ObjMesh(vertices=np.asarray(verts, np.float32).reshape(-1, 3),
                   colors=np.asarray(cols, np.float32).reshape(-1, 3), faces=faces,
                   texcoords=texcoords.astype(np.float32), texture=texture, mtl_path=mtl)


def read_mtl_texture(mtl_path: str):
    """The ``map_Kd`` image of an MTL file as (H, W, 3) float32 RGB in [0, 1], or None when the
    MTL names no image or the image file is absent."""
    from PIL import Image
    folder = os.path.dirname(mtl_path)
    tex = None
    with open(mtl_path) as fh:
        for line in fh:
            sp = line.split()
            if sp and sp[0] == "map_Kd" and os.path.exists(os.path.join(folder, sp[1])):
                with Image.open(os.path.join(folder, sp[1])) as im:
                    tex = np.asarray(im.convert("RGB"), np.float32) / 255.0
    return tex


@dataclass
class Calibration:
    extrinsics: np.ndarray        # (C, 12) float32, row-major [R|t]
    intrinsics: np.ndarray        # (C, 9) float32, row-major K at the render resolution

    @property
    def C(self):
        return self.extrinsics.shape[0]


def read_calibration(path: str, render_u: int, render_v: int) -> Calibration:
    extr, intr, size_u, size_v = [], [], [], []
    with open(path) as fh:
        for line in fh:
            sp = line.split()
            if not sp:
                continue
            if sp[0] == "intrinsic":
                intr += [float(sp[i]) for i in range(1, len(sp)) if i % 4 != 0 and i <= 11]
            elif sp[0] == "extrinsic":
                extr += [float(sp[i]) for i in range(1, len(sp)) if i <= 12]
            elif sp[0] == "size":
                size_u.append(float(sp[1]))
                size_v.append(float(sp[2]))
    C = len(extr) // 12
    K = np.asarray(intr, np.float64).reshape(C, 3, 3)
    for c in range(C):
        K[c, 0, 0] = K[c, 0, 0] / size_u[c] * render_u
        K[c, 1, 1] = K[c, 1, 1] / size_v[c] * render_v
        K[c, 0, 2] = K[c, 0, 2] / size_u[c] * render_u
        K[c, 1, 2] = K[c, 1, 2] / size_v[c] * render_v
    return Calibration(extrinsics=np.asarray(extr, np.float32).reshape(C, 12),
                       intrinsics=K.reshape(C, 9).astype(np.float32))
