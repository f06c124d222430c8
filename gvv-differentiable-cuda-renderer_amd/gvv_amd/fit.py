"""Analysis-by-synthesis fitting through the differentiable rasterizer.

The reference's own use of the renderer is a fit loop: render, compare with a target image,
back-propagate and step an optimizer (python/test_gradients_VertexColor.py:93-138,
test_gradients_SphericalHarmonics.py:95-130).  This module runs that loop at the scale of the
workload configuration it is measured on:

* a template mesh (capped cylinder, 175 x 200 grid + 2 poles = 35,002 vertices / 70,000 faces)
  with an 80-dimensional linear shape basis of smooth seeded displacement fields,
  orthonormalised and scaled to 0.02 m per unit coefficient;
* per subject: 80 shape coefficients, 27 SH coefficients (shared by its views) and a 6-DoF
  pose (axis-angle rotation about the template centre, translation in units of 0.1 m).  The
  pose is applied to the vertices before rendering: the reference returns zero gradients
  for extrinsics (python/CudaRenderer.py:215), so rigid motion has to live upstream;
* the template's vertex colours are U(0.2, 0.9) per vertex (a high-frequency albedo);
  FitConfig.colour_smoothing > 0 smooths them over the mesh for a low-frequency albedo;
* vertexColor + shaded rendering of 4 views at 512 x 512, masked L2 photometric loss against
  targets rendered from ground-truth parameters, Adam (lr 1e-2).  The loss and its gradient
  with respect to the render run on the library's own kernels (include/gvv_fit.h): one
  deterministic pass each over render, target and mask (MaskedSquaredError).

Lengths are millimetres inside the renderer (the reference's thresholds assume mm).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib, scenes
from .renderer import Rasterizer, RenderFunction, _ptr, _stream_ptr

MM = 1000.0


@dataclass
class FitConfig:
    n_basis: int = 80
    basis_scale_m: float = 0.02
    cols: int = 175
    rows: int = 200
    views: int = 4
    resolution: int = 512
    batch: int = 16
    iters: int = 200
    lr: float = 1e-2
    seed: int = 0
    colour_smoothing: int = 0     # Laplacian steps applied to the template's random vertex colours
    basis_smoothing: int = 60     # Laplacian steps applied to the Gaussian displacement fields (0: white)


def _smooth(faces, n, X, iters):
    """`iters` steps of uniform Laplacian averaging of per-vertex values X (n, m) over the face
    adjacency."""
    import scipy.sparse as sp
    rows = np.concatenate([faces[:, 0], faces[:, 1], faces[:, 2], faces[:, 1], faces[:, 2], faces[:, 0]])
    cols = np.concatenate([faces[:, 1], faces[:, 2], faces[:, 0], faces[:, 0], faces[:, 1], faces[:, 2]])
    A = sp.coo_matrix((np.ones(rows.size), (rows, cols)), shape=(n, n)).tocsr()
    A.data[:] = 1.0
    deg = np.asarray(A.sum(1)).ravel()
    M = sp.diags(1.0 / np.maximum(deg, 1)) @ A
    for _ in range(iters):
        X = 0.5 * X + 0.5 * (M @ X)
    return X


def _smooth_fields(rng, faces, n, k, iters=60):
    """k random 3-vector fields on the mesh vertices, smoothed by `iters` steps of uniform
    Laplacian averaging over the face adjacency (smooth, low-frequency displacements)."""
    X = _smooth(faces, n, rng.normal(size=(n, 3 * k)), iters)
    return X.reshape(n, k, 3).transpose(1, 0, 2)          # (k, n, 3)


def rodrigues(w: torch.Tensor) -> torch.Tensor:
    """(B, 3) axis-angle -> (B, 3, 3) rotation matrices, smooth at zero."""
    th2 = (w * w).sum(-1, keepdim=True)
    th = torch.sqrt(th2 + 1e-24)
    small = th2 < 1e-8
    a = torch.where(small, 1.0 - th2 / 6.0, torch.sin(th) / th)
    b = torch.where(small, 0.5 - th2 / 24.0, (1.0 - torch.cos(th)) / (th2 + 1e-24))
    zero = torch.zeros_like(w[:, 0])
    K = torch.stack([zero, -w[:, 2], w[:, 1], w[:, 2], zero, -w[:, 0], -w[:, 1], w[:, 0], zero], -1).view(-1, 3, 3)
    eye = torch.eye(3, device=w.device, dtype=w.dtype).expand(w.shape[0], 3, 3)
    return eye + a[..., None] * K + b[..., None] * (K @ K)


@dataclass
class Params:
    alpha: torch.Tensor        # (B, K) shape coefficients
    sh: torch.Tensor           # (B, 27)
    pose: torch.Tensor         # (B, 6): axis-angle (rad), translation (0.1 m units)

    def detach(self):
        return Params(self.alpha.detach().clone(), self.sh.detach().clone(), self.pose.detach().clone())

    def flat(self):
        return torch.cat([self.alpha, self.sh, self.pose], 1)

    @staticmethod
    def from_flat(x, n_basis):
        return Params(x[:, :n_basis], x[:, n_basis:n_basis + 27], x[:, n_basis + 27:n_basis + 33])


class ShapeModel:
    """Template, shape basis, colours and cameras of the fitting workload; renders parameters."""

    def __init__(self, cfg: FitConfig = FitConfig(), device=None):
        self.cfg = cfg
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        r_geo, r_basis, r_col = np.random.Generator(np.random.PCG64(cfg.seed)).spawn(3)
        v, f, uv = scenes.capped_cylinder(cfg.cols, cfg.rows, 1.0, 2.0, r_geo, 0.05)
        self.faces = f
        self.uv = uv
        self.N, self.F = v.shape[0], f.shape[0]
        B = _smooth_fields(r_basis, f, self.N, cfg.n_basis, cfg.basis_smoothing).reshape(cfg.n_basis, -1).T  # (3N, K)
        Q, _ = np.linalg.qr(B)
        Q = Q.T.reshape(cfg.n_basis, self.N, 3) * (cfg.basis_scale_m * MM * math.sqrt(self.N))
        dev = self.device
        self.template = torch.tensor(v * MM, dtype=torch.float32, device=dev)
        self.centre = self.template.mean(0)
        # (K, 3N), row-major: torch.tensor keeps a numpy array's strides, and Q is a transposed view
        self.basis = torch.tensor(np.ascontiguousarray(Q.reshape(cfg.n_basis, -1)), dtype=torch.float32, device=dev)
        col = r_col.uniform(0.2, 0.9, size=(self.N, 3))
        if cfg.colour_smoothing > 0:
            # low-frequency albedo: smoothed, then rescaled to the spread of U(0.2, 0.9) around 0.55
            c = _smooth(f, self.N, col - 0.55, cfg.colour_smoothing)
            col = np.clip(0.55 + c * (0.7 / np.sqrt(12.0) / max(c.std(), 1e-12)), 0.05, 0.95)
        self.colors = torch.tensor(col, dtype=torch.float32, device=dev)
        E, K = scenes.ring_cameras(cfg.views, 3.5, cfg.resolution, cfg.resolution)
        self.extr = torch.tensor(E, device=dev)
        self.intr = torch.tensor(K, device=dev)
        self.rast = Rasterizer(f.reshape(-1), uv.reshape(-1), self.N, cfg.views, cfg.resolution, cfg.resolution,
                               albedo_mode="vertexColor", shading_mode="shaded")
        self._tex = torch.full((1, 1, 1, 3), 0.5, device=dev)
        self._target_dummy = {}
        self._batched_cache = {}

    # -- geometry ------------------------------------------------------------------------------
    def motion(self, p: Params) -> torch.Tensor:
        """Per-subject rigid motion [B, 12]: rotation (row-major) and translation in mm."""
        B = p.pose.shape[0]
        R = rodrigues(p.pose[:, :3])
        t = p.pose[:, 3:6] * (0.1 * MM)
        return torch.cat([R.reshape(B, 9), t], 1)

    def vertices(self, p: Params) -> torch.Tensor:
        """Posed vertices [B, N, 3] on the shape-model kernels (ShapeVertices)."""
        return ShapeVertices.apply(p.alpha, p.pose, self.template, self.basis, self.centre)

    def vertices_reference(self, p: Params) -> torch.Tensor:
        """The same in torch operations (the kernels' reference)."""
        B = p.alpha.shape[0]
        X = self.template + (p.alpha @ self.basis).view(B, self.N, 3)
        m = self.motion(p)
        R, t = m[:, :9].view(B, 3, 3), m[:, 9:]
        return torch.matmul(X - self.centre, R.transpose(1, 2).contiguous()) + self.centre + t[:, None, :]

    def render(self, p: Params):
        """Returns (render [B,C,H,W,3], face ids [B,C,H,W]) for parameters p (differentiable)."""
        out = self.render_all(p)
        return out[2], out[1]

    def render_all(self, p: Params):
        """All six renderer outputs for parameters p (differentiable)."""
        B = p.alpha.shape[0]
        C, R = self.cfg.views, self.cfg.resolution
        vpos = self.vertices(p)
        sh = p.sh[:, None, :].expand(B, C, 27)
        key = B
        if key not in self._target_dummy:
            self._target_dummy[key] = torch.zeros((B, C, R, R, 3), device=self.device)
        return RenderFunction.apply(self.rast, vpos, self._batched(self.colors, B, self.N, 3),
                                    self._batched(self._tex, B, 1, 1, 3), sh, self._target_dummy[key],
                                    self._batched(self.extr, B, C, 12).view(B, C * 12),
                                    self._batched(self.intr, B, C, 9).view(B, C * 9))

    def _batched(self, t, B, *shape):
        """The per-subject copy [B, *shape] of a tensor shared by all subjects (colours, texture,
        cameras), made contiguous once and kept while `t` is the same, unmodified tensor: the
        renderer takes contiguous batches, and expanding them every iteration costs two copies
        per tensor (forward and backward)."""
        cache = self._batched_cache
        hit = cache.get((id(t), B))
        if hit is None or hit[0] is not t or hit[1] != t._version:
            hit = (t, t._version, t.expand(B, *shape).contiguous())
            cache[(id(t), B)] = hit
        return hit[2]


def sample_params(rng: np.random.Generator, B: int, n_basis: int, device) -> Params:
    """Ground-truth prior: shape ~ N(0, 0.5^2) per coefficient, SH DC ~ U(0.6, 1.0) with the
    other coefficients ~ N(0, 0.1^2), rotation ~ N(0, 0.1^2) rad per axis, translation
    ~ N(0, 0.3^2) x 0.1 m per axis."""
    alpha = rng.normal(0.0, 0.5, size=(B, n_basis))
    sh = rng.normal(0.0, 0.1, size=(B, 27))
    sh[:, [0, 9, 18]] = rng.uniform(0.6, 1.0, size=(B, 3))
    pose = np.concatenate([rng.normal(0.0, 0.1, size=(B, 3)), rng.normal(0.0, 0.3, size=(B, 3))], 1)
    t = lambda a: torch.tensor(a, dtype=torch.float32, device=device)
    return Params(t(alpha), t(sh), t(pose))


def plain_init(B: int, n_basis: int, device) -> Params:
    """The plain fit's start: zero shape and pose, neutral white DC lighting."""
    sh = torch.zeros((B, 27), device=device)
    sh[:, [0, 9, 18]] = 0.8
    return Params(torch.zeros((B, n_basis), device=device), sh, torch.zeros((B, 6), device=device))


def plain_subject_losses(model: "ShapeModel", target, mask):
    """Per-subject masked loss [B] of the plain start.  The plain start is the same for every
    subject (zero shape and pose, neutral lighting), so one subject's render serves all."""
    with torch.no_grad():
        r = model.render(plain_init(1, model.cfg.n_basis, model.device))[0]
        d = (r - target) * mask
        return (d * d).flatten(1).sum(1) / mask.flatten(1).sum(1).clamp_min(1.0)


class ShapeVertices(torch.autograd.Function):
    """v[b] = R_b (template + alpha[b] @ basis - centre) + centre + t_b on gvv_shape_forward, with
    R_b = rodrigues(pose[b, :3]) and t_b = pose[b, 3:6] x 0.1 m, and the gradients of alpha and
    pose from gvv_shape_backward.  Template [N, 3], basis [K, 3N] and centre [3] take no
    gradient."""

    TSCALE = 0.1 * MM     # translation units of the pose, in mm

    @staticmethod
    def forward(ctx, alpha, pose, template, basis, centre):
        B, K = alpha.shape
        N = template.shape[0]
        if pose.shape != (B, 6) or template.shape != (N, 3) or basis.shape != (K, 3 * N) or centre.numel() != 3:
            raise ValueError(f"shape vertices: alpha {tuple(alpha.shape)}, pose {tuple(pose.shape)}, template "
                             f"{tuple(template.shape)}, basis {tuple(basis.shape)}, centre {tuple(centre.shape)}")
        for name, t in (("alpha", alpha), ("pose", pose), ("template", template), ("basis", basis),
                        ("centre", centre)):
            if t.dtype != torch.float32 or not t.is_cuda:
                raise ValueError(f"shape vertices: {name} must be a float32 device tensor")
        alpha, pose, template, basis, centre = (t.contiguous() for t in (alpha, pose, template, basis, centre))
        dev = alpha.device
        v = torch.empty((B, N, 3), device=dev)
        xc = torch.empty((B, N, 3), device=dev)
        rt = torch.empty((B, 12), device=dev)
        L = _lib.load()
        _lib.check(L.gvv_shape_forward(B, N, K, _ptr(alpha), _ptr(pose), ShapeVertices.TSCALE, _ptr(template),
                                       _ptr(basis), _ptr(centre), _ptr(v), _ptr(xc), _ptr(rt), _stream_ptr(dev)),
                   "gvv_shape_forward")
        ctx.save_for_backward(xc, pose, rt, basis)
        return v

    @staticmethod
    def backward(ctx, g_v):
        xc, pose, rt, basis = ctx.saved_tensors
        B, N, _ = xc.shape
        K = basis.shape[0]
        g_v = g_v.contiguous()
        L = _lib.load()
        work = torch.empty(int(L.gvv_shape_workspace_floats(B, N, K)), device=xc.device)
        g = torch.empty((B, K + 6), device=xc.device)
        _lib.check(L.gvv_shape_backward(B, N, K, _ptr(g_v), _ptr(xc), _ptr(pose), _ptr(rt), ShapeVertices.TSCALE,
                                        _ptr(basis), _ptr(work), _ptr(g), _stream_ptr(xc.device)),
                   "gvv_shape_backward")
        return g[:, :K], g[:, K:], None, None, None


class MaskedSquaredError(torch.autograd.Function):
    """Per-subject masked squared error sq[b] = sum over pixels and channels of
    ((render - target) * mask)^2, for render/target [B, ..., 3] and mask [B, ..., 1] (the same
    pixels), on gvv_masked_sq_error; the backward (gvv_masked_sq_error_grad) gives the render
    gradient 2 (render - target) mask^2 g[b].  Target and mask take no gradient."""

    @staticmethod
    def forward(ctx, render, target, mask):
        B = render.shape[0]
        if target.shape != render.shape or render.shape[-1] != 3 or mask.shape != render.shape[:-1] + (1,):
            raise ValueError(f"masked squared error: render {tuple(render.shape)}, target {tuple(target.shape)} and "
                             f"mask {tuple(mask.shape)} must be [B, ..., 3], [B, ..., 3] and [B, ..., 1]")
        for name, t in (("render", render), ("target", target), ("mask", mask)):
            if t.dtype != torch.float32 or not t.is_cuda:
                raise ValueError(f"masked squared error: {name} must be a float32 device tensor")
        render, target, mask = render.contiguous(), target.contiguous(), mask.contiguous()
        P = mask.numel() // B
        L = _lib.load()
        dev = render.device
        work = torch.empty(int(L.gvv_masked_sq_workspace_floats(B, P)), device=dev)
        sq = torch.empty(B, device=dev)
        _lib.check(L.gvv_masked_sq_error(B, P, _ptr(render), _ptr(target), _ptr(mask), _ptr(work), _ptr(sq),
                                         _stream_ptr(dev)), "gvv_masked_sq_error")
        ctx.save_for_backward(render, target, mask)
        return sq

    @staticmethod
    def backward(ctx, g_sq):
        render, target, mask = ctx.saved_tensors
        B = render.shape[0]
        P = mask.numel() // B
        g_sq = g_sq.contiguous()
        g_render = torch.empty_like(render)
        L = _lib.load()
        _lib.check(L.gvv_masked_sq_error_grad(B, P, _ptr(render), _ptr(target), _ptr(mask), _ptr(g_sq),
                                              _ptr(g_render), _stream_ptr(render.device)), "gvv_masked_sq_error_grad")
        return g_render, None, None


def masked_sq_error(render, target, mask):
    """Per-subject masked squared error [B] (MaskedSquaredError)."""
    return MaskedSquaredError.apply(render, target, mask)


def masked_l2(render, target, mask):
    """Mean squared photometric error over the target's foreground pixels, per subject summed."""
    d = (render - target) * mask
    return (d * d).sum() / mask.sum().clamp_min(1.0)


@dataclass
class FitResult:
    params: Params
    losses: list = field(default_factory=list)
    times: list = field(default_factory=list)          # wall-clock seconds when each loss was known
    fallbacks: int = 0

    @property
    def final_loss(self):
        return self.losses[-1]

    def time_to(self, loss_level):
        """Wall-clock seconds until the loss first reached loss_level (None if never)."""
        for l, t in zip(self.losses, self.times):
            if l <= loss_level:
                return t
        return None


def fit(model: ShapeModel, target, mask, init: Params, iters=None, lr=None, stop_at=None, t0=None,
        fallback=None) -> FitResult:
    """Adam over (shape, SH, pose).  `losses[i]` is the loss of the parameters before step i+1;
    `times[i]` the wall-clock since t0 (default: entry) once that loss is known on the host.
    Stops as soon as a loss is <= stop_at, returning the parameters that reached it.

    The host reads each iteration's loss one iteration late: iteration i+1's forward is queued
    before the host waits for loss i (an event after its copy to pinned memory), so the GPU does
    not idle while the host queues work.  Stopping at loss i therefore drops the forward already
    queued for iteration i+1 and returns the parameters of iteration i (kept on the device).

    `fallback` = (plain parameters, their per-subject losses [B]): the first iteration's forward
    also yields each subject's own loss at `init` (read at once, not late); a subject whose loss
    there is worse than its plain loss falls back to the plain parameters (its first update is
    skipped, as its gradient was taken at the discarded start) and the recorded first loss is
    that of the chosen starts.  `FitResult.fallbacks` counts them."""
    cfg = model.cfg
    iters = cfg.iters if iters is None else iters
    lr = cfg.lr if lr is None else lr
    x = init.flat().detach().clone().requires_grad_(True)
    opt = torch.optim.Adam([x], lr=lr)
    res = FitResult(params=None)
    t0 = time.perf_counter() if t0 is None else t0
    mask_sub = mask.flatten(1).sum(1)
    msum = mask_sub.sum().clamp_min(1.0)
    host = torch.empty(2, dtype=torch.float32, pin_memory=True)     # in-flight losses, by iteration parity
    snaps = [torch.empty_like(x.detach()) for _ in range(2)]          # the parameters they belong to
    pending = None                                                    # (slot, event) of the unread loss
    done = None                                                       # parameters that reached stop_at

    def read(slot, ev):
        ev.synchronize()
        lv = float(host[slot])
        res.losses.append(lv)
        res.times.append(time.perf_counter() - t0)
        return stop_at is not None and lv <= stop_at

    for it in range(iters):
        slot = it & 1
        with torch.no_grad():
            snaps[slot].copy_(x)
        opt.zero_grad(set_to_none=True)
        render, _ = model.render(Params.from_flat(x, cfg.n_basis))
        sq = masked_sq_error(render, target, mask)           # per subject
        loss = sq.sum() / msum
        back = None
        if it == 0 and fallback is not None:
            plain, plain_losses = fallback
            own = sq.detach() / mask_sub.clamp_min(1.0)
            back = own > plain_losses
            res.fallbacks = int(back.sum())
            if res.fallbacks:
                chosen = torch.where(back, plain_losses * mask_sub, sq.detach())
                lv = float(chosen.sum() / msum)
                with torch.no_grad():
                    snaps[slot][back] = plain.flat()[back]
            else:
                lv = float(loss.detach())
            res.losses.append(lv)
            res.times.append(time.perf_counter() - t0)
            if stop_at is not None and lv <= stop_at:
                done = snaps[slot]
                break
        else:
            host[slot].copy_(loss.detach(), non_blocking=True)
            ev = torch.cuda.Event()
            ev.record()
            if pending is not None and read(*pending):
                done = snaps[slot ^ 1]
                pending = None
                break
            pending = (slot, ev)
        loss.backward()
        if back is not None and res.fallbacks:
            with torch.no_grad():
                x.grad[back] = 0.0
                x[back] = plain.flat()[back]
        opt.step()
    if pending is not None and read(*pending):
        done = snaps[pending[0]]
    res.params = Params.from_flat((x if done is None else done).detach().clone(), cfg.n_basis)
    return res


def make_targets(model: ShapeModel, gt: Params):
    with torch.no_grad():
        render, face = model.render(gt)
    mask = (face >= 0).to(torch.float32).unsqueeze(-1)
    return render.contiguous(), mask
