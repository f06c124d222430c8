"""Seeded synthetic scenes for the renderer workloads.

Geometry is generated in metres and handed to the renderer in millimetres (x1000 on vertex
positions and camera translations; intrinsics stay in pixels).  The reference's absolute
thresholds assume millimetres: rayTriangleIntersect rejects |d.N|<1e-7 after scaling vertices by
1/1000 (cpp/src/Utils/RendererUtil.h:29-47), so a unit-scale mesh renders nothing.

Cameras follow the reference convention: extrinsics are the row-major 3x4 world->camera [R|t]
(cpp/src/Renderer/CUDABasedRasterization.cu:37-48), camera +z looks forward and +y points down,
and intrinsics are the row-major 3x3 K.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

MM = 1000.0


@dataclass
class Scene:
    faces: np.ndarray        # (F, 3) int32
    uv: np.ndarray           # (F, 3, 2) float32, per-face per-corner (u, v)
    vpos: np.ndarray         # (B, N, 3) float32, millimetres
    vcol: np.ndarray         # (B, N, 3) float32
    tex: np.ndarray          # (B, texH, texW, 3) float32
    sh: np.ndarray           # (B, C, 27) float32
    extr: np.ndarray         # (B, C, 12) float32
    intr: np.ndarray         # (B, C, 9) float32
    target: np.ndarray       # (B, C, H, W, 3) float32
    width: int
    height: int
    albedo_mode: str = "vertexColor"
    shading_mode: str = "shaded"
    meta: dict = field(default_factory=dict)

    @property
    def B(self):
        return self.vpos.shape[0]

    @property
    def C(self):
        return self.extr.shape[1]

    @property
    def N(self):
        return self.vpos.shape[1]

    @property
    def F(self):
        return self.faces.shape[0]

    def renderer_kwargs(self):
        return dict(faces=self.faces.reshape(-1), texcoords=self.uv.reshape(-1), n_vertices=self.N,
                    n_cameras=self.C, width=self.width, height=self.height,
                    albedo_mode=self.albedo_mode, shading_mode=self.shading_mode)

    def inputs(self):
        return self.vpos, self.vcol, self.tex, self.sh, self.extr, self.intr


def _rngs(seed, n):
    return np.random.Generator(np.random.PCG64(seed)).spawn(n)


# ---------------------------------------------------------------------------------------------
# cameras
# ---------------------------------------------------------------------------------------------
def look_at(position_m, target_m=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)):
    """World->camera [R|t] in millimetres for a camera at ``position_m`` looking at ``target_m``."""
    c = np.asarray(position_m, np.float64) * MM
    f = np.asarray(target_m, np.float64) * MM - c
    f /= np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, np.float64))
    if np.linalg.norm(r) < 1e-6:
        r = np.cross(f, np.array([0.0, 0.0, 1.0]))
    r /= np.linalg.norm(r)
    d = np.cross(f, r)
    R = np.stack([r, d, f])
    t = -R @ c
    return np.concatenate([R, t[:, None]], axis=1).astype(np.float32)


def pinhole(width, height, focal_scale=1.2):
    f = focal_scale * width
    return np.array([[f, 0.0, width / 2.0], [0.0, f, height / 2.0], [0.0, 0.0, 1.0]], np.float32)


def ring_cameras(n, radius_m, width, height, y_m=0.0):
    E = [look_at((radius_m * np.sin(2 * np.pi * k / n), y_m, radius_m * np.cos(2 * np.pi * k / n)))
         for k in range(n)]
    K = [pinhole(width, height)] * n
    return np.stack(E).reshape(n, 12), np.stack(K).reshape(n, 9)


def sphere_cameras(n, radius_m, width, height, rng):
    E = []
    while len(E) < n:
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        if abs(d[1]) > 0.9:
            continue
        E.append(look_at(radius_m * d))
    K = [pinhole(width, height)] * n
    return np.stack(E).reshape(n, 12), np.stack(K).reshape(n, 9)


def sh_coefficients(n_cameras, rng, dc=0.8, sigma=0.1):
    sh = rng.normal(0.0, sigma, size=27).astype(np.float32)
    sh[[0, 9, 18]] = dc
    return np.tile(sh[None], (n_cameras, 1)).astype(np.float32)


# ---------------------------------------------------------------------------------------------
# meshes
# ---------------------------------------------------------------------------------------------
def icosphere(level):
    """Subdivided icosahedron on the unit sphere: 10*4^level+2 vertices, 20*4^level faces."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    faces = f
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.asarray(verts, np.float64), np.asarray(faces, np.int32)


def _smooth_noise_field(rng, n_terms=6, max_freq=3):
    """Sum of random low-frequency sinusoids in (theta, y); returns f(theta, y) in [-1, 1]."""
    kt = rng.integers(1, max_freq + 1, size=n_terms)
    ky = rng.uniform(0.5, max_freq, size=n_terms)
    ph = rng.uniform(0, 2 * np.pi, size=(n_terms, 2))
    amp = rng.uniform(0.5, 1.0, size=n_terms)
    amp /= amp.sum()

    def f(theta, y):
        s = np.zeros(np.broadcast(theta, y).shape)
        for i in range(n_terms):
            s += amp[i] * np.sin(kt[i] * theta + ph[i, 0]) * np.cos(ky[i] * np.pi * y + ph[i, 1])
        return s

    return f


def capped_cylinder(cols, rows, radius_m, height_m, rng, amplitude_m):
    """Grid of ``rows`` x ``cols`` vertices wrapped around the y axis plus two pole vertices.

    Faces: (rows-1)*cols*2 quads-as-triangles + 2*cols pole fans; 250x200 gives 50,002 vertices
    and 100,000 faces.  Per-face UVs come from the grid parameterisation (u around, v up).
    """
    noise = _smooth_noise_field(rng)
    theta = 2 * np.pi * np.arange(cols) / cols
    ys = np.linspace(-height_m / 2, height_m / 2, rows)
    T, Y = np.meshgrid(theta, ys)                          # (rows, cols)
    R = radius_m + amplitude_m * noise(T, Y / (height_m / 2))
    grid = np.stack([R * np.sin(T), Y, R * np.cos(T)], -1).reshape(-1, 3)
    south = np.array([[0.0, -height_m / 2, 0.0]])
    north = np.array([[0.0, height_m / 2, 0.0]])
    verts = np.concatenate([grid, south, north], 0)
    S, Nn = rows * cols, rows * cols + 1

    def vid(r, c):
        return r * cols + (c % cols)

    faces, uvs = [], []
    for r in range(rows - 1):
        v0, v1 = r / (rows - 1), (r + 1) / (rows - 1)
        for c in range(cols):
            u0, u1 = c / cols, (c + 1) / cols
            a, b, cc, d = vid(r, c), vid(r, c + 1), vid(r + 1, c + 1), vid(r + 1, c)
            faces += [(a, b, cc), (a, cc, d)]
            uvs += [((u0, v0), (u1, v0), (u1, v1)), ((u0, v0), (u1, v1), (u0, v1))]
    for c in range(cols):
        u0, u1, um = c / cols, (c + 1) / cols, (c + 0.5) / cols
        faces.append((S, vid(0, c + 1), vid(0, c)))
        uvs.append(((um, 0.0), (u1, 0.0), (u0, 0.0)))
        faces.append((Nn, vid(rows - 1, c), vid(rows - 1, c + 1)))
        uvs.append(((um, 1.0), (u0, 1.0), (u1, 1.0)))
    return verts, np.asarray(faces, np.int32), np.asarray(uvs, np.float32)


def smooth_texture(rng, height, width, coarse=16):
    """Bilinearly upsampled random grid: smooth noise in [0, 1]."""
    g = rng.uniform(0.0, 1.0, size=(coarse + 1, coarse + 1, 3))
    ys = np.linspace(0, coarse, height, endpoint=False) + 0.5 * coarse / height
    xs = np.linspace(0, coarse, width, endpoint=False) + 0.5 * coarse / width
    y0 = np.clip(np.floor(ys).astype(int), 0, coarse - 1)
    x0 = np.clip(np.floor(xs).astype(int), 0, coarse - 1)
    fy = (ys - y0)[:, None, None]
    fx = (xs - x0)[None, :, None]
    a = g[y0][:, x0]
    b = g[y0][:, x0 + 1]
    c = g[y0 + 1][:, x0]
    d = g[y0 + 1][:, x0 + 1]
    return ((1 - fy) * ((1 - fx) * a + fx * b) + fy * ((1 - fx) * c + fx * d)).astype(np.float32)


# ---------------------------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------------------------
def config0(seed=0, width=256, height=256, level=4, n_cameras=4):
    """Icosphere (5,120 faces / 2,562 vertices at level 4), radius 1 m jittered by N(0, 1 cm),
    vertex colours U(0,1), SH with DC 0.8 and N(0,0.1) others, 4 cameras on a 3 m sphere,
    f = 1.2 W, vertexColor + shaded."""
    r_geo, r_col, r_sh, r_cam, r_tgt = _rngs(seed, 5)
    v, f = icosphere(level)
    v = v + r_geo.normal(0.0, 0.01, size=v.shape)
    N = v.shape[0]
    vpos = (v * MM).astype(np.float32)[None]
    vcol = r_col.uniform(0.0, 1.0, size=(1, N, 3)).astype(np.float32)
    E, K = sphere_cameras(n_cameras, 3.0, width, height, r_cam)
    sh = sh_coefficients(n_cameras, r_sh)[None]
    target = r_tgt.uniform(0.0, 1.0, size=(1, n_cameras, height, width, 3)).astype(np.float32)
    return Scene(faces=f, uv=np.zeros((f.shape[0], 3, 2), np.float32), vpos=vpos, vcol=vcol,
                 tex=np.full((1, 1, 1, 3), 0.5, np.float32), sh=sh, extr=E[None], intr=K[None],
                 target=target, width=width, height=height, albedo_mode="vertexColor",
                 shading_mode="shaded", meta=dict(name="icosphere5120_4x256", seed=seed))


def config1(seed=0, width=1024, height=1024, cols=250, rows=200, n_cameras=8, tex_size=1024,
            albedo_mode="textured", with_target=True):
    """Capped cylinder 250x200 + 2 poles (50,002 vertices / 100,000 faces), radius 1 m, height
    2 m, radial noise amplitude 5 cm, UVs from the grid, 1024^2 smooth-noise texture, 8 views on
    a 3.5 m ring at 1024^2, textured + shaded."""
    r_geo, r_col, r_sh, r_tex, r_tgt = _rngs(seed, 5)
    v, f, uv = capped_cylinder(cols, rows, 1.0, 2.0, r_geo, 0.05)
    N = v.shape[0]
    vpos = (v * MM).astype(np.float32)[None]
    vcol = r_col.uniform(0.0, 1.0, size=(1, N, 3)).astype(np.float32)
    E, K = ring_cameras(n_cameras, 3.5, width, height)
    sh = sh_coefficients(n_cameras, r_sh)[None]
    tex = smooth_texture(r_tex, tex_size, tex_size)[None]
    if with_target:
        target = r_tgt.uniform(0.0, 1.0, size=(1, n_cameras, height, width, 3)).astype(np.float32)
    else:
        target = np.zeros((1, n_cameras, height, width, 3), np.float32)
    return Scene(faces=f, uv=uv, vpos=vpos, vcol=vcol, tex=tex, sh=sh, extr=E[None], intr=K[None],
                 target=target, width=width, height=height, albedo_mode=albedo_mode,
                 shading_mode="shaded",
                 meta=dict(name=f"cylinder{f.shape[0]}_{n_cameras}x{width}", seed=seed))


def config2(seed=0, n_meshes=32, tri_range=(2000, 200000), resolutions=(256, 512, 1024), max_views=8):
    """Batched uneven workload: n_meshes meshes with triangle counts log-uniform in tri_range,
    alternately icospheres (the subdivision level whose count is nearest in log scale, within
    the range) and capped-cylinder grids (cols x rows chosen for the drawn count); each with
    1..max_views views (uniform) at one square resolution drawn from `resolutions`, its own
    vertex colours, SH and cameras on a 3 m sphere as in config0.  vertexColor + shaded."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    lo, hi = np.log(tri_range[0]), np.log(tri_range[1])
    levels = [L for L in range(1, 9) if tri_range[0] <= 20 * 4 ** L <= tri_range[1]]
    for k in range(n_meshes):
        r_geo, r_col, r_sh, r_cam, r_pick = rng.spawn(5)
        F_target = float(np.exp(r_pick.uniform(lo, hi)))
        C = int(r_pick.integers(1, max_views + 1))
        R = int(resolutions[r_pick.integers(0, len(resolutions))])
        if k % 2 == 0 and levels:
            L = min(levels, key=lambda L: abs(np.log(20 * 4 ** L) - np.log(F_target)))
            v, f = icosphere(L)
            v = v + r_geo.normal(0.0, 0.01, size=v.shape)
            uv = np.zeros((f.shape[0], 3, 2), np.float32)
            kind = f"icosphere{f.shape[0]}"
        else:
            cols = max(3, int(round(np.sqrt(F_target * 1.25 / 2.0))))
            rows = max(2, int(round(F_target / (2.0 * cols))))
            v, f, uv = capped_cylinder(cols, rows, 0.8, 1.6, r_geo, 0.04)
            v = v - v.mean(0)
            kind = f"grid{f.shape[0]}"
        N = v.shape[0]
        E, K = sphere_cameras(C, 3.0, R, R, r_cam)
        out.append(Scene(faces=f, uv=uv, vpos=(v * MM).astype(np.float32)[None],
                         vcol=r_col.uniform(0.0, 1.0, size=(1, N, 3)).astype(np.float32),
                         tex=np.full((1, 1, 1, 3), 0.5, np.float32), sh=sh_coefficients(C, r_sh)[None],
                         extr=E[None], intr=K[None], target=np.zeros((1, C, R, R, 3), np.float32), width=R, height=R,
                         albedo_mode="vertexColor", shading_mode="shaded",
                         meta=dict(name=f"{kind}_{C}x{R}", seed=seed, index=k)))
    return out


def single_triangle(width=64, height=64, z_m=2.0, albedo_mode="vertexColor", shading_mode="shaded",
                    tri_m=((-0.5, -0.4, 0.0), (0.5, -0.4, 0.0), (0.0, 0.6, 0.0))):
    """One triangle in the plane z=0 seen by a camera on the -z axis looking at +z."""
    v = np.asarray(tri_m, np.float64)
    E = look_at((0.0, 0.0, -z_m), up=(0.0, 1.0, 0.0))
    K = pinhole(width, height, 1.0)
    sh = np.zeros(27, np.float32)
    sh[[0, 9, 18]] = 1.0
    uv = np.array([[[0.1, 0.1], [0.9, 0.1], [0.5, 0.9]]], np.float32)
    return Scene(faces=np.array([[0, 1, 2]], np.int32), uv=uv, vpos=(v * MM).astype(np.float32)[None],
                 vcol=np.array([[[1, 0, 0], [0, 1, 0], [0, 0, 1]]], np.float32),
                 tex=np.linspace(0, 1, 8 * 8 * 3, dtype=np.float32).reshape(1, 8, 8, 3),
                 sh=sh[None, None], extr=E.reshape(1, 1, 12), intr=K.reshape(1, 1, 9),
                 target=np.zeros((1, 1, height, width, 3), np.float32), width=width, height=height,
                 albedo_mode=albedo_mode, shading_mode=shading_mode, meta=dict(name="triangle"))


def random_soup(seed, n_faces=40, n_cameras=2, width=48, height=40, albedo_mode="vertexColor",
                shading_mode="shaded", tex_size=(16, 12), batch=1):
    """Small random mesh: a jittered icosphere-like blob plus free triangles, for brute-force and
    finite-difference checks.  Every vertex belongs to at least one face."""
    rng = np.random.Generator(np.random.PCG64(seed))
    v, f = icosphere(1)
    v = v * 0.6 + rng.normal(0, 0.03, size=v.shape)
    extra = []
    for _ in range(max(0, n_faces - f.shape[0])):
        c = rng.uniform(-0.8, 0.8, size=3)
        extra.append(c + rng.normal(0, 0.25, size=(3, 3)))
    if extra:
        ev = np.concatenate(extra, 0)
        ef = np.arange(ev.shape[0], dtype=np.int32).reshape(-1, 3) + v.shape[0]
        v = np.concatenate([v, ev], 0)
        f = np.concatenate([f, ef], 0)
    N, F = v.shape[0], f.shape[0]
    vpos = np.stack([(v * MM + rng.normal(0, 1.0, size=v.shape)).astype(np.float32) for _ in range(batch)])
    vcol = rng.uniform(0, 1, size=(batch, N, 3)).astype(np.float32)
    E, K = sphere_cameras(n_cameras, 3.0, width, height, rng)
    sh = np.stack([sh_coefficients(n_cameras, rng) for _ in range(batch)])
    th, tw = tex_size
    tex = rng.uniform(0, 1, size=(batch, th, tw, 3)).astype(np.float32)
    uv = rng.uniform(0.05, 0.95, size=(F, 3, 2)).astype(np.float32)
    target = rng.uniform(0, 1, size=(batch, n_cameras, height, width, 3)).astype(np.float32)
    return Scene(faces=f.astype(np.int32), uv=uv, vpos=vpos, vcol=vcol, tex=tex, sh=sh,
                 extr=np.tile(E[None], (batch, 1, 1)), intr=np.tile(K[None], (batch, 1, 1)),
                 target=target, width=width, height=height, albedo_mode=albedo_mode,
                 shading_mode=shading_mode, meta=dict(name=f"soup{seed}"))


def stress_soup(seed, n_faces=80, width=64, height=48, albedo_mode="vertexColor"):
    """Hard cases for visibility: triangles crossing the camera plane or behind the camera,
    triangles covering most of the image, edge-on and sub-pixel triangles, coincident faces."""
    rng = np.random.Generator(np.random.PCG64(seed))
    E = look_at((0.0, 0.0, -3.0))
    K = pinhole(width, height, 1.0)
    tris = []
    for i in range(n_faces):
        kind = i % 5
        if kind == 0:      # crossing / behind the camera plane (camera at z = -3 m)
            c = np.array([rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), rng.uniform(-3.5, -2.0)])
            tris.append(c + rng.normal(0, 0.8, size=(3, 3)))
        elif kind == 1:    # large, close
            c = np.array([0.0, 0.0, rng.uniform(-2.0, 0.5)])
            tris.append(c + rng.normal(0, 1.5, size=(3, 3)) * np.array([1, 1, 0.1]))
        elif kind == 2:    # edge-on: all three vertices nearly in a plane through the camera ray
            c = rng.uniform(-0.5, 0.5, size=3)
            dz = rng.normal(0, 0.4)
            tris.append(np.array([c, c + np.array([1e-4, 0, dz]), c + np.array([-1e-4, 1e-4, -dz])]))
        elif kind == 3:    # sub-pixel
            c = rng.uniform(-0.6, 0.6, size=3)
            tris.append(c + rng.normal(0, 2e-3, size=(3, 3)))
        else:              # ordinary
            c = rng.uniform(-0.6, 0.6, size=3)
            tris.append(c + rng.normal(0, 0.2, size=(3, 3)))
    v = np.concatenate(tris, 0)
    f = np.arange(v.shape[0], dtype=np.int32).reshape(-1, 3)
    f = np.concatenate([f, f[:4]], 0)                      # coincident duplicates
    N, F = v.shape[0], f.shape[0]
    sh = sh_coefficients(1, rng)
    return Scene(faces=f, uv=rng.uniform(0, 1, size=(F, 3, 2)).astype(np.float32),
                 vpos=(v * MM).astype(np.float32)[None], vcol=rng.uniform(0, 1, size=(1, N, 3)).astype(np.float32),
                 tex=rng.uniform(0, 1, size=(1, 16, 16, 3)).astype(np.float32), sh=sh[None],
                 extr=E.reshape(1, 1, 12), intr=K.reshape(1, 1, 9),
                 target=rng.uniform(0, 1, size=(1, 1, height, width, 3)).astype(np.float32),
                 width=width, height=height, albedo_mode=albedo_mode, meta=dict(name=f"stress{seed}"))


def behind_camera(with_near=True, width=48, height=48):
    """A triangle with one vertex 0.1 m behind the camera, in the plane 1 m off the optical axis,
    its other two vertices 4 m ahead and outside the image on both sides; with `with_near`, also
    a large triangle 0.01 mm in front of the camera covering the whole image (face 1).  The
    geometry is given in camera coordinates (mm) and mapped to world coordinates."""
    E = look_at((0.0, 0.0, -2.0))
    K = pinhole(width, height, 1.0)
    R, t = E[:, :3].astype(np.float64), E[:, 3].astype(np.float64)
    world = lambda cam_pts: (R.T @ (np.asarray(cam_pts, np.float64) - t).T).T
    tris = [world([[0.0, 1000.0, -100.0], [-3000.0, 1000.0, 4000.0], [3000.0, 1000.0, 4000.0]])]
    if with_near:
        tris.append(world([[-10.0, -10.0, 0.01], [10.0, -10.0, 0.01], [0.0, 10.0, 0.01]]))
    v = np.concatenate(tris).astype(np.float32)
    n = len(tris)
    rng = np.random.Generator(np.random.PCG64(3))
    return Scene(faces=np.arange(3 * n, dtype=np.int32).reshape(n, 3),
                 uv=rng.uniform(0, 1, size=(n, 3, 2)).astype(np.float32), vpos=v[None],
                 vcol=rng.uniform(0, 1, size=(1, 3 * n, 3)).astype(np.float32),
                 tex=rng.uniform(0, 1, size=(1, 8, 8, 3)).astype(np.float32), sh=sh_coefficients(1, rng)[None],
                 extr=E.reshape(1, 1, 12), intr=K.reshape(1, 1, 9),
                 target=rng.uniform(0, 1, size=(1, 1, height, width, 3)).astype(np.float32),
                 width=width, height=height, meta=dict(name="behind_camera"))


# SH preset of the reference's demo scripts (python/data/test_SH_tensor.py:4): per channel
# 0.7 ambient and -0.5 on the x band.
REFERENCE_SH_PRESET = np.array([0.7, 0, 0, -0.5, 0, 0, 0, 0, 0] * 3, np.float32)


def from_fixture(mesh, calib, width, height, batch=1, albedo_mode="vertexColor", shading_mode="shaded",
                 texture=None, sh=None, seed=0):
    """Scene from an OBJ mesh (gvv_amd.io.read_obj) and a calibration (gvv_amd.io.read_calibration),
    with the reference demo's SH preset unless `sh` is given."""
    rng = np.random.Generator(np.random.PCG64(seed))
    C = calib.C
    tex = texture if texture is not None else mesh.texture
    if tex is None:
        tex = smooth_texture(rng, 64, 64)
    sh = REFERENCE_SH_PRESET if sh is None else np.asarray(sh, np.float32)
    return Scene(faces=mesh.faces, uv=mesh.texcoords, vpos=np.tile(mesh.vertices[None], (batch, 1, 1)),
                 vcol=np.tile(mesh.colors[None], (batch, 1, 1)),
                 tex=np.tile(tex.astype(np.float32)[None], (batch, 1, 1, 1)),
                 sh=np.tile(sh.reshape(1, 1, 27), (batch, C, 1)).astype(np.float32),
                 extr=np.tile(calib.extrinsics[None], (batch, 1, 1)), intr=np.tile(calib.intrinsics[None], (batch, 1, 1)),
                 target=rng.uniform(0, 1, size=(batch, C, height, width, 3)).astype(np.float32),
                 width=width, height=height, albedo_mode=albedo_mode, shading_mode=shading_mode,
                 meta=dict(name=os.path.basename(mesh.mtl_path or "mesh")))
