"""Pins of the CPU oracle against what the mathematics fixes, independent of the oracle's own
formulas: closed forms, brute force, invariants and central finite differences.

Each test is chosen so that a plausible mistake (a dropped term, a wrong sign or index, a
transposed operand) fails it.  The reference ships no golden vectors or known-answer tests
for this path (its python/test_*.py scripts are interactive demos), so these are the pins."""
import numpy as np
import pytest

from conftest import rel_err
from gvv_amd import scenes
from oracle import oracle as O


def _camera_matrices(scene, c=0):
    E = scene.extr[0, c].reshape(3, 4).astype(np.float64)
    K = scene.intr[0, c].reshape(3, 3).astype(np.float64)
    return E, K


# ----------------------------------------------------------------------------------------------
# camera set-up
# ----------------------------------------------------------------------------------------------
def test_camera_inverse_identities():
    rng = np.random.Generator(np.random.PCG64(0))
    for _ in range(5):
        pos = rng.normal(size=3)
        pos *= 3.0 / np.linalg.norm(pos)
        E = scenes.look_at(pos)
        K = scenes.pinhole(640, 480, 1.1)
        invE, invP, o = O.camera(E.reshape(-1), K.reshape(-1))
        R, t = E[:, :3].astype(np.float64), E[:, 3].astype(np.float64)
        # rigid inverse in closed form: [R^T, -R^T t]
        assert np.allclose(invE[:3, :3], R.T, atol=1e-6)
        assert np.allclose(invE[:3, 3], -R.T @ t, rtol=1e-6, atol=1e-2)
        assert np.allclose(invE[3], [0, 0, 0, 1], atol=1e-6)
        # the ray origin is the camera centre
        assert np.allclose(o, pos * 1000.0, rtol=1e-5, atol=1e-2)
        K4, E4 = np.eye(4), np.eye(4)
        K4[:3, :3], E4[:3, :] = K, E
        inv64 = np.linalg.inv(K4 @ E4)
        assert np.allclose(invP, inv64, rtol=1e-4, atol=1e-6 * np.abs(inv64).max())


# ----------------------------------------------------------------------------------------------
# visibility
# ----------------------------------------------------------------------------------------------
def _analytic_plane_bary(scene, px, py):
    """FP64 barycentrics of the hit of pixel (px, py) on the z=0 triangle of single_triangle."""
    E, K = _camera_matrices(scene)
    R, t = E[:, :3], E[:, 3]
    c = -R.T @ t
    d = R.T @ np.linalg.solve(K, np.array([px, py, 1.0]))
    lam = -c[2] / d[2]
    X = c + lam * d
    v = scene.vpos[0].astype(np.float64)
    T = np.array([[v[0, 0], v[1, 0], v[2, 0]], [v[0, 1], v[1, 1], v[2, 1]], [1, 1, 1]])
    return np.linalg.solve(T, np.array([X[0], X[1], 1.0]))


def test_single_triangle_closed_form():
    s = scenes.single_triangle(width=64, height=64)
    orc = O.Oracle(**s.renderer_kwargs())
    out = orc.forward(*s.inputs())
    face, bary = out["face"][0, 0], out["bary"][0, 0]
    checked = 0
    for py in range(64):
        for px in range(64):
            abc = _analytic_plane_bary(s, px + 0.5, py + 0.5)
            if abc.min() > 1e-4:
                assert face[py, px] == 0
                assert abs(bary[py, px, 0] - abc[0]) < 1e-5 and abs(bary[py, px, 1] - abc[1]) < 1e-5
                checked += 1
            elif abc.min() < -2e-3:
                assert face[py, px] == -1
    assert checked > 300


def test_equal_depth_tie_lowest_id():
    """Faces that repeat the same vertices in the same order produce bit-identical depth keys at
    every pixel; the lowest face id among them must win (the reference leaves this to an
    unsynchronised last write)."""
    s = scenes.single_triangle(width=32, height=32)
    other = np.array([[-0.9, -0.9, 0.5], [-0.6, -0.9, 0.5], [-0.75, -0.6, 0.5]], np.float32) * 1000.0
    vpos = np.concatenate([s.vpos[0], other])[None]
    for faces, winner in (([[0, 1, 2], [0, 1, 2], [0, 1, 2]], 0), ([[3, 4, 5], [0, 1, 2], [0, 1, 2]], 1)):
        sc = scenes.Scene(**{**s.__dict__})
        sc.vpos, sc.faces = vpos, np.array(faces, np.int32)
        sc.vcol = np.ones((1, 6, 3), np.float32)
        sc.uv = np.zeros((3, 3, 2), np.float32)
        out = O.Oracle(**sc.renderer_kwargs()).forward(*sc.inputs())
        f = out["face"][0, 0]
        covered = f[f >= 0]
        assert covered.size > 100
        assert set(np.unique(covered).tolist()) <= {winner, 0 if winner == 1 else winner}
        assert (f == winner).sum() > 100


def test_nearer_triangle_wins():
    W = H = 24
    E = scenes.look_at((0.0, 0.0, -2.0))
    K = scenes.pinhole(W, H, 1.0)
    big = np.array([[-0.5, -0.5, 0.0], [0.5, -0.5, 0.0], [0.0, 0.6, 0.0]])
    near = big * 0.5 + np.array([0.0, 0.0, -0.2])          # closer to the camera
    v = (np.concatenate([big, near]) * 1000.0).astype(np.float32)
    for order in ([[0, 1, 2], [3, 4, 5]], [[3, 4, 5], [0, 1, 2]]):
        faces = np.array(order, np.int32)
        sc = scenes.Scene(faces=faces, uv=np.zeros((2, 3, 2), np.float32), vpos=v[None],
                          vcol=np.ones((1, 6, 3), np.float32), tex=np.ones((1, 1, 1, 3), np.float32),
                          sh=np.zeros((1, 1, 27), np.float32), extr=E.reshape(1, 1, 12), intr=K.reshape(1, 1, 9),
                          target=np.zeros((1, 1, H, W, 3), np.float32), width=W, height=H)
        out = O.Oracle(**sc.renderer_kwargs()).forward(*sc.inputs())
        near_id = 1 if order[1][0] == 3 else 0
        assert out["face"][0, 0, H // 2, W // 2] == near_id


def _ray_hits(E, K, tri, px, py):
    """FP64 barycentrics of the hit of pixel (px, py)'s ray on triangle tri, or None when the
    ray misses its plane in front of the camera."""
    R, t = E[:, :3], E[:, 3]
    c = -R.T @ t
    d = R.T @ np.linalg.solve(K, np.array([px, py, 1.0]))
    n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
    den = d @ n
    if abs(den) < 1e-12:
        return None
    lam = ((tri[0] - c) @ n) / den
    if lam <= 0:
        return None
    X = c + lam * d
    T = np.stack([tri[0] - tri[2], tri[1] - tri[2]], axis=1)
    ab = np.linalg.lstsq(T, X - tri[2], rcond=None)[0]
    return np.array([ab[0], ab[1], 1.0 - ab[0] - ab[1]])


def test_vertex_behind_the_camera():
    """A triangle with one vertex 0.1 m behind the camera (the reference clamps that vertex's
    camera depth to 1e-5 before the perspective divide, the numerators K c keeping the unclamped
    depth, CameraUtil.h:157-169).  Two consequences that follow from the geometry and fail for
    plausible mistakes:
    (a) coverage: the clamped vertex projects to (f x + cx z, f y + cy z) / 1e-5, far out
    towards the image's bottom-left here, so the bounding box reaches the image border on that
    side and the covered pixels are exactly the pixels whose rays hit the triangle in front of
    the camera (with the depth used unclamped, negative, the vertex lands above the image and
    the box keeps one row of them);
    (b) depth: the clamped vertex's inverse depth 1e5 dominates the interpolated depth, so the
    triangle wins over a triangle 0.01 mm in front of the camera, although that one is nearer
    along every ray (a clamp to 1e-2 instead would make the near triangle win)."""
    both = scenes.behind_camera()
    W, H = both.width, both.height
    E64, K64 = _camera_matrices(both)
    v = both.vpos[0].astype(np.float64)
    tri, near = v[:3], v[3:]

    def render(n):
        sc = scenes.behind_camera(with_near=n == 2)
        return O.Oracle(**sc.renderer_kwargs()).forward(*sc.inputs())["face"][0, 0]

    f = render(1)
    inside = 0
    for py in range(H):
        for px in range(W):
            abc = _ray_hits(E64, K64, tri, px + 0.5, py + 0.5)
            if abc is not None and abc.min() > 1e-6:
                assert f[py, px] == 0, (px, py, abc)
                inside += 1
            elif abc is None or abc.min() < -1e-3:
                assert f[py, px] == -1, (px, py)
    assert inside > 300

    f2 = render(2)
    won = lost = 0
    for py in range(H):
        for px in range(W):
            t_abc = _ray_hits(E64, K64, tri, px + 0.5, py + 0.5)
            n_abc = _ray_hits(E64, K64, near, px + 0.5, py + 0.5)
            assert n_abc is not None and n_abc.min() > 1e-3
            if t_abc is not None and t_abc.min() > 1e-6 and t_abc[0] > 1.5e-4:
                assert f2[py, px] == 0, (px, py, t_abc)
                won += 1
            elif t_abc is None or t_abc.min() < -1e-3:
                assert f2[py, px] == 1, (px, py)
                lost += 1
    assert won > 300 and lost > 300


def _brute_force_ids(scene):
    """Every triangle at every pixel, smallest (depth key, id); independent of bounding boxes."""
    B, C = scene.B, scene.C
    H, W = scene.height, scene.width
    out = np.full((B, C, H, W), -1, np.int32)
    for b in range(B):
        for c in range(C):
            E = scene.extr[b, c].astype(np.float32)
            K = scene.intr[b, c].astype(np.float32)
            Em, Km = E.reshape(3, 4).astype(np.float32), K.reshape(3, 3).astype(np.float32)
            best = np.full((H, W), np.iinfo(np.uint64).max, np.uint64)
            for f, tri in enumerate(scene.faces):
                v = scene.vpos[b, tri].astype(np.float32)
                # projected camera z of each vertex (float32, reference operand order)
                zs = []
                for p in v:
                    cam = [np.float32(Em[r, 0] * p[0]) + np.float32(Em[r, 1] * p[1]) for r in range(3)]
                    cam = [np.float32(np.float32(cam[r] + np.float32(Em[r, 2] * p[2])) + Em[r, 3]) for r in range(3)]
                    z = np.float32(np.float32(cam[0] * Km[2, 0]) + np.float32(cam[1] * Km[2, 1]))
                    z = np.float32(z + np.float32(cam[2] * Km[2, 2]))
                    zs.append(z if z > np.float32(1e-7) else np.float32(1e-5))
                for py in range(H):
                    for px in range(W):
                        _, abc = O.pixel_barycentric(E, K, px + 0.5, py + 0.5, v.reshape(-1))
                        if not ((abc >= -0.001).all() and (abc <= 1.001).all()):
                            continue
                        with np.errstate(all="ignore"):
                            s = np.float32(np.float32(abc[0] / zs[0]) + np.float32(abc[1] / zs[1]))
                            s = np.float32(s + np.float32(abc[2] / zs[2]))
                            z = np.float32(np.float32(np.float32(1.0) / s) * np.float32(10000.0))
                        zq = int(np.clip(np.nan_to_num(z, nan=0.0), -2**31, 2**31 - 1))
                        key = np.uint64(((zq & 0xffffffff) ^ 0x80000000) << 32 | f)
                        if key < best[py, px]:
                            best[py, px] = key
            hit = best != np.iinfo(np.uint64).max
            out[b, c][hit] = (best[hit] & np.uint64(0xffffffff)).astype(np.int32)
    return out


def test_brute_force_visibility_small_soup():
    s = scenes.random_soup(11, n_faces=60, n_cameras=2, width=20, height=16)
    out = O.Oracle(**s.renderer_kwargs()).forward(*s.inputs())
    ref = _brute_force_ids(s)
    assert (ref >= 0).sum() > 100
    assert np.array_equal(out["face"], ref)


def test_visibility_invariants():
    s = scenes.random_soup(3, n_faces=100, n_cameras=2, width=48, height=40)
    faces_by_mode = []
    for albedo, shading in (("vertexColor", "shaded"), ("textured", "shadeless"), ("normal", "shaded"),
                            ("lighting", "shaded"), ("foregroundMask", "shaded")):
        kw = s.renderer_kwargs()
        kw.update(albedo_mode=albedo, shading_mode=shading)
        out = O.Oracle(**kw).forward(*s.inputs())
        faces_by_mode.append(out["face"])
        bg = out["face"] < 0
        assert np.all(out["render"][bg] == np.array([0.0, 1.0, 0.0]))
        assert np.all(out["bary"][bg] == 0.0)
        a, b = out["bary"][..., 0][~bg], out["bary"][..., 1][~bg]
        assert (a >= 0).all() and (b >= 0).all() and (a + b <= 1.001 + 1e-6).all()
    for f in faces_by_mode[1:]:
        assert np.array_equal(f, faces_by_mode[0])


# ----------------------------------------------------------------------------------------------
# shading
# ----------------------------------------------------------------------------------------------
def test_sh_basis_on_plane_facing_camera():
    """A plane facing the camera has its (flipped) normal along -z of the world here, so the
    lighting mode renders sh0 - sh2 + 2 sh6 per channel and the normal mode (0.5, 0.5, 0)."""
    s = scenes.single_triangle()
    rng = np.random.Generator(np.random.PCG64(5))
    s.sh = rng.normal(size=(1, 1, 27)).astype(np.float32)
    kw = s.renderer_kwargs()
    kw["albedo_mode"] = "lighting"
    out = O.Oracle(**kw).forward(*s.inputs())
    m = out["face"][0, 0] >= 0
    sh = s.sh[0, 0].astype(np.float64).reshape(3, 9)
    expect = sh[:, 0] - sh[:, 2] + 2.0 * sh[:, 6]
    assert np.allclose(out["render"][0, 0][m], expect, rtol=1e-12, atol=1e-12)
    kw["albedo_mode"] = "normal"
    out = O.Oracle(**kw).forward(*s.inputs())
    assert np.allclose(out["render"][0, 0][m], [0.5, 0.5, 0.0], atol=1e-12)


def test_nearest_texel_on_ramp():
    s = scenes.single_triangle(albedo_mode="textured", shading_mode="shadeless")
    th, tw = 8, 8
    ramp = np.zeros((1, th, tw, 3), np.float32)
    ramp[0, :, :, 0] = np.arange(tw)[None, :]          # column index
    ramp[0, :, :, 1] = np.arange(th)[:, None]          # row index
    s.tex = ramp
    out = O.Oracle(**s.renderer_kwargs()).forward(*s.inputs())
    m = out["face"][0, 0] >= 0
    a, b = out["bary"][0, 0][..., 0][m], out["bary"][0, 0][..., 1][m]
    c = np.float32(1.0) - a - b
    uv = s.uv[0]
    u = (np.float32(uv[0, 0]) * a + np.float32(uv[1, 0]) * b) + np.float32(uv[2, 0]) * c
    v = ((np.float32(1) - uv[0, 1]) * a + (np.float32(1) - uv[1, 1]) * b) + (np.float32(1) - uv[2, 1]) * c
    U = np.clip(u * np.float32(tw), 0, tw - 1)
    V = np.clip(v * np.float32(th), 0, th - 1)
    col = np.trunc(U - np.float32(0.5)).astype(int)
    row = np.trunc(V - np.float32(0.5)).astype(int)
    assert np.array_equal(out["render"][0, 0][m][:, 0], col.astype(np.float64))
    assert np.array_equal(out["render"][0, 0][m][:, 1], row.astype(np.float64))


# ----------------------------------------------------------------------------------------------
# topology
# ----------------------------------------------------------------------------------------------
def test_reference_vertex_faces_isolated_vertex():
    faces = np.array([[0, 2, 3], [2, 3, 0]], np.int32)          # vertex 1 is isolated
    lst, pairs = O.reference_vertex_faces(faces, 4)
    assert lst.tolist() == [0, 1, 0, 1, 0, 1]
    # the isolated vertex gets no (start, count) entry, so entries after it are shifted
    assert pairs.tolist() == [[0, 2], [2, 2], [4, 2]]


# ----------------------------------------------------------------------------------------------
# gradients: central finite differences of the oracle's own forward
# ----------------------------------------------------------------------------------------------
def _loss(orc, scene, w, vpos=None, vcol=None, tex=None, sh=None, mask=None):
    out = orc.forward(scene.vpos if vpos is None else vpos, scene.vcol if vcol is None else vcol,
                      scene.tex if tex is None else tex, scene.sh if sh is None else sh, scene.extr, scene.intr)
    r = out["render"]
    if mask is not None:
        r = np.where(mask[..., None], r, 0.0)
    return float((w * r).sum()), out


def _front_facing(scene, out, margin=2e-2):
    """Covered pixels whose interpolated normal faces the camera by a margin: the reference
    flips the normal elsewhere and then skips texture gradients and uses an unflipped
    Jacobian, so finite differences only match on these pixels."""
    m = np.zeros(out["face"].shape, bool)
    for c in range(scene.C):
        E, K = _camera_matrices(scene, c)
        R, t = E[:, :3], E[:, 3]
        cen = -R.T @ t
        f = out["face"][0, c]
        ys, xs = np.nonzero(f >= 0)
        tri = scene.faces[f[ys, xs]]
        a = out["bary"][0, c, ys, xs, 0].astype(np.float64)
        b = out["bary"][0, c, ys, xs, 1].astype(np.float64)
        vn = out["vnormal"][0, c].astype(np.float64)
        n = a[:, None] * vn[tri[:, 0]] + b[:, None] * vn[tri[:, 1]] + (1 - a - b)[:, None] * vn[tri[:, 2]]
        n /= np.linalg.norm(n, axis=1, keepdims=True)
        d = (R.T @ np.linalg.solve(K, np.stack([xs + 0.5, ys + 0.5, np.ones_like(xs, float)]))).T
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        ok = (n * d).sum(1) < -margin
        m[0, c, ys[ok], xs[ok]] = True
    return m


def _grads(orc, scene, out, w):
    return orc.backward(w.astype(np.float32), np.zeros_like(w, np.float32), scene.target, scene.vpos, scene.vcol,
                        scene.tex, scene.sh, out["vnormal"], out["bary"], out["face"], scene.extr, scene.intr)


@pytest.fixture(scope="module")
def sphere_scene():
    return scenes.config0(width=96, height=96, level=2, n_cameras=2)


def test_fd_vertex_color_and_sh(sphere_scene):
    s = sphere_scene
    orc = O.Oracle(**s.renderer_kwargs())
    w = np.random.Generator(np.random.PCG64(1)).normal(size=(1, s.C, s.height, s.width, 3))
    _, out = _loss(orc, s, w)
    g = _grads(orc, s, out, w)
    h = 1e-2
    rng = np.random.Generator(np.random.PCG64(2))
    for v in rng.choice(s.N, 6, replace=False):
        for k in range(3):
            cp, cm = s.vcol.copy(), s.vcol.copy()
            cp[0, v, k] += h
            cm[0, v, k] -= h
            fd = (_loss(orc, s, w, vcol=cp)[0] - _loss(orc, s, w, vcol=cm)[0]) / float(cp[0, v, k] - cm[0, v, k])
            assert abs(fd - g["col"][0, v, k]) <= 1e-6 * max(1.0, abs(fd))
    for c in range(s.C):
        for k in range(27):
            sp, sm = s.sh.copy(), s.sh.copy()
            sp[0, c, k] += h
            sm[0, c, k] -= h
            fd = (_loss(orc, s, w, sh=sp)[0] - _loss(orc, s, w, sh=sm)[0]) / float(sp[0, c, k] - sm[0, c, k])
            assert abs(fd - g["sh"][0, c, k]) <= 1e-6 * max(1.0, abs(fd))


def test_fd_texture(sphere_scene):
    s = sphere_scene
    s = scenes.Scene(**{**s.__dict__})
    rng = np.random.Generator(np.random.PCG64(3))
    s.tex = rng.uniform(0, 1, size=(1, 12, 10, 3)).astype(np.float32)
    s.uv = rng.uniform(0, 1, size=s.uv.shape).astype(np.float32)
    s.albedo_mode = "textured"
    orc = O.Oracle(**s.renderer_kwargs())
    w = rng.normal(size=(1, s.C, s.height, s.width, 3))
    _, out = _loss(orc, s, w)
    w = np.where(_front_facing(s, out)[..., None], w, 0.0)
    g = _grads(orc, s, out, w)
    h = 1e-2
    for (r, c) in [(0, 0), (3, 4), (7, 9), (11, 2)]:
        for k in range(3):
            tp, tm = s.tex.copy(), s.tex.copy()
            tp[0, r, c, k] += h
            tm[0, r, c, k] -= h
            fd = (_loss(orc, s, w, tex=tp)[0] - _loss(orc, s, w, tex=tm)[0]) / float(tp[0, r, c, k] - tm[0, r, c, k])
            assert abs(fd - g["tex"][0, r, c, k]) <= 1e-6 * max(1.0, abs(fd))


def test_fd_vertex_position_through_normals(sphere_scene):
    """Constant albedo makes the render depend on positions only through the interpolated
    normal, which the reference differentiates exactly (barycentric and adjacency terms).
    Pixels whose face id changes under the perturbation are excluded from both sides."""
    s = scenes.Scene(**{**sphere_scene.__dict__})
    s.vcol = np.full_like(s.vcol, 0.7)
    orc = O.Oracle(**s.renderer_kwargs())
    rng = np.random.Generator(np.random.PCG64(4))
    w = rng.normal(size=(1, s.C, s.height, s.width, 3))
    _, base = _loss(orc, s, w)
    w = np.where(_front_facing(s, base)[..., None], w, 0.0)
    h = 0.25                                            # millimetres; edges are ~150 mm
    gfull = _grads(orc, s, base, w)["pos"][0]
    scale = np.abs(gfull).max()
    verts = np.argsort(-np.abs(gfull).max(1))[:6]
    for v in verts:
        for k in range(3):
            vp, vm = s.vpos.copy(), s.vpos.copy()
            vp[0, v, k] += h
            vm[0, v, k] -= h
            op = orc.forward(vp, s.vcol, s.tex, s.sh, s.extr, s.intr)
            om = orc.forward(vm, s.vcol, s.tex, s.sh, s.extr, s.intr)
            stable = (op["face"] == base["face"]) & (om["face"] == base["face"]) & (base["face"] >= 0)
            wm = np.where(stable[..., None], w, 0.0)
            fd = float((wm * (op["render"] - om["render"])).sum()) / float(vp[0, v, k] - vm[0, v, k])
            g = _grads(orc, s, base, wm)["pos"][0, v, k]
            assert abs(fd - g) <= 1e-2 * scale, (v, k, fd, g, scale)


def test_model_to_data_term_on_linear_ramp(sphere_scene):
    """With a linear-ramp target I(x, y) = alpha x + beta y the image-gradient filter returns
    exactly (alpha, beta), so the model-to-data gradient is the derivative of
    sum_p gT_p . I(project(sum_r bcc_r v_r)) with the barycentrics held fixed."""
    s = sphere_scene
    alpha, beta = np.array([0.3, -0.2, 0.1]), np.array([-0.05, 0.4, 0.25])
    ys, xs = np.mgrid[0:s.height, 0:s.width]
    ramp = xs[..., None] * alpha + ys[..., None] * beta
    target = np.broadcast_to(ramp, (1, s.C, s.height, s.width, 3)).astype(np.float32).copy()
    orc = O.Oracle(**s.renderer_kwargs())
    out = orc.forward(*s.inputs())
    rng = np.random.Generator(np.random.PCG64(6))
    gT = rng.normal(size=target.shape).astype(np.float32)
    fs = 1
    border = np.zeros((s.height, s.width), bool)
    border[fs + 1:s.height - (fs + 1), fs + 1:s.width - (fs + 1)] = True
    g = orc.backward(np.zeros_like(gT), gT, target, s.vpos, s.vcol, s.tex, s.sh, out["vnormal"], out["bary"],
                     out["face"], s.extr, s.intr)["pos"][0]

    def energy(vpos):
        e = 0.0
        for c in range(s.C):
            E, K = _camera_matrices(s, c)
            f = out["face"][0, c]
            m = (f >= 0) & border
            a = out["bary"][0, c][..., 0][m].astype(np.float64)
            b = out["bary"][0, c][..., 1][m].astype(np.float64)
            tri = s.faces[f[m]]
            X = a[:, None] * vpos[tri[:, 0]] + b[:, None] * vpos[tri[:, 1]] + (1 - a - b)[:, None] * vpos[tri[:, 2]]
            P = (K @ (E[:, :3] @ X.T + E[:, 3:4])).T
            uvp = P[:, :2] / P[:, 2:3]
            I = uvp[:, :1] * alpha + uvp[:, 1:2] * beta
            e += float((gT[0, c][m] * I).sum())
        return e

    V = s.vpos[0].astype(np.float64)
    h = 1e-3
    for v in rng.choice(s.N, 6, replace=False):
        for k in range(3):
            Vp, Vm = V.copy(), V.copy()
            Vp[v, k] += h
            Vm[v, k] -= h
            fd = (energy(Vp) - energy(Vm)) / (2 * h)
            assert abs(fd - g[v, k]) <= 1e-6 * max(1.0, abs(fd)), (v, k, fd, g[v, k])


def test_bary_jacobian_finite_difference():
    """d(a,b,c)/d(vertices) with the ray fixed, against central differences of the FP64
    ray/triangle barycentrics (closed-form plane intersection)."""
    rng = np.random.Generator(np.random.PCG64(8))
    for _ in range(5):
        tri = rng.normal(size=(3, 3)) * 100.0
        centre = tri.mean(0)
        o = centre + rng.normal(size=3) * 50 + np.array([0, 0, -3000.0])
        w = rng.dirichlet([2, 2, 2])
        d = (w @ tri) - o
        d /= np.linalg.norm(d)
        tri32, o32, d32 = tri.astype(np.float32), o.astype(np.float32), d.astype(np.float32)
        J = O.bary_jacobian(o32, d32, tri32.reshape(-1))

        def bary(t):
            n = np.cross(t[1] - t[0], t[2] - t[0])
            tt = (t[0] - o32) @ n / (d32 @ n)
            P = o32 + tt * d32
            a = n @ np.cross(t[2] - t[1], P - t[1]) / (n @ n)
            b = n @ np.cross(t[0] - t[2], P - t[2]) / (n @ n)
            return np.array([a, b, 1 - a - b])

        base = tri32.astype(np.float64)
        for var in range(9):
            tp, tm = base.copy(), base.copy()
            tp[var // 3, var % 3] += 1e-3
            tm[var // 3, var % 3] -= 1e-3
            fd = (bary(tp) - bary(tm)) / 2e-3
            assert np.allclose(J[:, var], fd, rtol=1e-5, atol=1e-9)


# ----------------------------------------------------------------------------------------------
# further pins: normals, interpolation, normal map, textured-backward quirks, grazing exit
# ----------------------------------------------------------------------------------------------
def test_vertex_normals_are_incident_face_sums():
    s = scenes.config0(width=32, height=32, level=2, n_cameras=1)
    out = O.Oracle(**s.renderer_kwargs()).forward(*s.inputs())
    v = s.vpos[0].astype(np.float64)
    fn = np.cross(v[s.faces[:, 1]] - v[s.faces[:, 0]], v[s.faces[:, 2]] - v[s.faces[:, 0]])
    vn = np.zeros_like(v)
    for k in range(3):
        np.add.at(vn, s.faces[:, k], fn)
    got = out["vnormal"][0, 0].astype(np.float64)
    assert np.allclose(got, vn, rtol=1e-5, atol=1e-3 * np.abs(vn).max())
    assert ((got * v).sum(1) > 0).all()              # outward on the convex sphere


def test_vertex_color_interpolation_identity_colours():
    """Red/green/blue corners under unit lighting render exactly (a, b, 1-a-b)."""
    s = scenes.single_triangle()                     # sh = DC 1 per channel, colours = identity
    out = O.Oracle(**s.renderer_kwargs()).forward(*s.inputs())
    m = out["face"][0, 0] >= 0
    a = out["bary"][0, 0][..., 0][m].astype(np.float64)
    b = out["bary"][0, 0][..., 1][m].astype(np.float64)
    assert np.allclose(out["render"][0, 0][m], np.stack([a, b, 1 - a - b], 1), atol=1e-12)


def test_normal_map_single_triangle():
    """Texels inside the triangle's UV footprint get (n+1)/2 of its unit normal, the rest 0.5."""
    s = scenes.single_triangle(albedo_mode="normal")
    s.tex = np.zeros((1, 40, 32, 3), np.float32)
    kw = s.renderer_kwargs()
    out = O.Oracle(**kw, compute_normal_map=True).forward(*s.inputs())
    nm = out["normal_map"][0]
    assert (out["face"] == -1).all()                 # rasterisation is skipped in this mode
    n = np.array([0.0, 0.0, 1.0])                    # face normal of the z=0 triangle
    uv = s.uv[0].astype(np.float64) * np.array([32, 40]) * np.array([1, -1]) + np.array([0, 40])
    T = np.array([[uv[0, 0], uv[1, 0], uv[2, 0]], [uv[0, 1], uv[1, 1], uv[2, 1]], [1, 1, 1]])
    inside = outside = 0
    for y in range(40):
        for x in range(32):
            w = np.linalg.solve(T, np.array([x + 0.5, y + 0.5, 1.0]))
            if w.min() > 1e-3:
                assert np.allclose(nm[y, x], (n + 1) / 2, atol=1e-12)
                inside += 1
            elif w.min() < -1e-3:
                assert np.allclose(nm[y, x], 0.5, atol=1e-12)
                outside += 1
    assert inside > 100 and outside > 100


def test_textured_backward_albedo_is_reflected_blend():
    """Grad.cu:311-312 blends the four texels with swapped weights: on a ramp c = col index the
    backward albedo is LU_index + 1 - s where U0 = LU_index + 0.5 + s (a standard bilinear blend
    would give LU_index + s).  Read through the DC SH gradient of single pixels."""
    s = scenes.single_triangle(albedo_mode="textured")
    # reverse the winding so pixels are not flipped; normal (0,0,-1) faces the camera
    s.faces = s.faces[:, ::-1].copy()
    s.uv = s.uv[:, ::-1].copy()
    th, tw = 8, 8
    ramp = np.zeros((1, th, tw, 3), np.float32)
    ramp[0, :, :, 0] = np.arange(tw)[None, :]
    s.tex = ramp
    orc = O.Oracle(**s.renderer_kwargs())
    out = orc.forward(*s.inputs())
    ys, xs = np.nonzero(out["face"][0, 0] >= 0)
    checked = 0
    for y, x in list(zip(ys, xs))[::37]:
        g = np.zeros((1, 1, s.height, s.width, 3), np.float32)
        g[0, 0, y, x, 0] = 1.0
        gr = orc.backward(g, np.zeros_like(g), s.target, s.vpos, s.vcol, s.tex, s.sh, out["vnormal"], out["bary"],
                          out["face"], s.extr, s.intr)
        a, b = out["bary"][0, 0, y, x]
        c = np.float32(1.0) - a - b
        uv = s.uv[0]
        u = (np.float32(uv[0, 0]) * a + np.float32(uv[1, 0]) * b) + np.float32(uv[2, 0]) * c
        U = min(max(np.float32(u * np.float32(tw)), np.float32(0)), np.float32(tw - 1))
        lu = int(np.trunc(U - np.float32(0.5)))
        sfrac = float(U) - (lu + 0.5)
        if lu + 1 >= tw:
            continue
        expect = lu + 1 - sfrac                          # x SH basis Y_0 = 1
        assert abs(gr["sh"][0, 0, 0] - expect) < 1e-9, (y, x, gr["sh"][0, 0, 0], expect)
        checked += 1
    assert checked > 5


def test_texture_gradient_skipped_for_flipped_pixels():
    """Grad.cu:345: pixels whose interpolated normal faces away receive no texture gradient."""
    s = scenes.single_triangle(albedo_mode="textured")       # winding gives +z: every pixel flipped
    g = np.ones((1, 1, s.height, s.width, 3), np.float32)
    orc = O.Oracle(**s.renderer_kwargs())
    out = orc.forward(*s.inputs())
    gr = orc.backward(g, np.zeros_like(g), s.target, s.vpos, s.vcol, s.tex, s.sh, out["vnormal"], out["bary"],
                      out["face"], s.extr, s.intr)
    assert (out["face"] >= 0).sum() > 100 and np.all(gr["tex"] == 0)
    s.faces = s.faces[:, ::-1].copy()                        # reversed winding: nothing flipped
    orc = O.Oracle(**s.renderer_kwargs())
    out = orc.forward(*s.inputs())
    gr = orc.backward(g, np.zeros_like(g), s.target, s.vpos, s.vcol, s.tex, s.sh, out["vnormal"], out["bary"],
                      out["face"], s.extr, s.intr)
    assert np.abs(gr["tex"]).sum() > 0


def test_bary_jacobian_grazing_exit():
    """|cos(ray, N)| < 0.001 returns a zero Jacobian (RendererUtil.h:682)."""
    tri = np.array([[0, 0, 0], [100, 0, 0], [0, 100, 0]], np.float32)
    o = np.array([30, 30, -3000], np.float32)
    for cz, zero in ((5e-4, True), (5e-2, False)):
        d = np.array([1.0, 0.0, cz], np.float64)
        d /= np.linalg.norm(d)
        J = O.bary_jacobian(o, d.astype(np.float32), tri.reshape(-1))
        assert (np.abs(J).max() == 0.0) == zero


def test_zero_gradients_in_normal_and_lighting_modes():
    s = scenes.random_soup(2, n_faces=40, width=24, height=24)
    for mode in ("normal", "lighting"):
        kw = s.renderer_kwargs()
        kw["albedo_mode"] = mode
        orc = O.Oracle(**kw)
        out = orc.forward(*s.inputs())
        g = np.ones(out["render"].shape, np.float32)
        gr = orc.backward(g, g, s.target, s.vpos, s.vcol, s.tex, s.sh, out["vnormal"], out["bary"], out["face"],
                          s.extr, s.intr)
        assert all(np.all(v == 0) for v in gr.values())
