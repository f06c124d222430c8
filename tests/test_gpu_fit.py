"""GPU checks of the analysis-by-synthesis fit harness (gvv_amd.fit): the autograd chain
params -> vertices -> renderer -> masked L2 against central finite differences (h = 1e-3),
the renderer's vertex gradients on the fit scene against the FP64 oracle, and a short fit."""
import numpy as np
import pytest
import torch

from conftest import rel_err
from gvv_amd import fit
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def small_cfg():
    return fit.FitConfig(cols=60, rows=50, views=2, resolution=128, batch=2, n_basis=8)


def front_facing(model, out, margin=0.05):
    """Covered pixels whose interpolated normal faces the camera by `margin`."""
    bary, face, vnormal = out[0], out[1], out[3]
    B, C, H, W = face.shape
    faces = torch.as_tensor(model.faces, device=face.device).long()
    f = face.clamp_min(0).long()
    tri = faces[f]                                                     # B,C,H,W,3
    vn = torch.stack([torch.gather(vnormal[:, c].unsqueeze(1).expand(B, H * W, model.N, 3), 2,
                                   tri[:, c].reshape(B, H * W, 3, 1).expand(B, H * W, 3, 3)) for c in range(C)], 1)
    vn = vn.view(B, C, H, W, 3, 3)
    a, b = bary[..., 0], bary[..., 1]
    w = torch.stack([a, b, 1 - a - b], -1)
    n = (w[..., None] * vn).sum(-2)
    n = n / n.norm(dim=-1, keepdim=True).clamp_min(1e-30)
    E = model.extr.view(C, 3, 4).double()
    K = model.intr.view(C, 3, 3).double()
    ys, xs = torch.meshgrid(torch.arange(H, device=face.device), torch.arange(W, device=face.device), indexing="ij")
    pix = torch.stack([xs + 0.5, ys + 0.5, torch.ones_like(xs, dtype=torch.float64)], -1).double()
    d = torch.einsum("cji,chwj->chwi", E[:, :, :3], torch.einsum("cij,hwj->chwi", torch.linalg.inv(K), pix))
    d = d / d.norm(dim=-1, keepdim=True)
    cos = (n.double() * d[None]).sum(-1)
    return (face >= 0) & (cos < -margin)


def test_fd_parameter_gradients(device):
    cfg = small_cfg()
    model = fit.ShapeModel(cfg, device)
    model.colors = torch.full_like(model.colors, 0.6)       # albedo independent of barycentrics
    rng = np.random.Generator(np.random.PCG64(3))
    gt = fit.sample_params(rng, cfg.batch, cfg.n_basis, device)
    target, mask = fit.make_targets(model, gt)
    p0 = fit.sample_params(np.random.Generator(np.random.PCG64(4)), cfg.batch, cfg.n_basis, device)
    with torch.no_grad():
        base = model.render_all(p0)
    keep = front_facing(model, base)

    def loss_of(x, stable):
        r, _ = model.render(fit.Params.from_flat(x, cfg.n_basis))
        m = mask * stable.unsqueeze(-1)
        d = (r.double() - target.double()) * m
        return (d * d).sum()

    x = p0.flat().detach().clone()
    h = 1e-3
    idx = [(0, 0), (0, 3), (1, 5), (0, cfg.n_basis + 0), (1, cfg.n_basis + 13), (0, cfg.n_basis + 27),
           (1, cfg.n_basis + 29), (0, cfg.n_basis + 31)]
    worst = 0.0
    for (b, k) in idx:
        xp, xm = x.clone(), x.clone()
        xp[b, k] += h
        xm[b, k] -= h
        with torch.no_grad():
            op = model.render_all(fit.Params.from_flat(xp, cfg.n_basis))
            om = model.render_all(fit.Params.from_flat(xm, cfg.n_basis))
        stable = keep & (op[1] == base[1]) & (om[1] == base[1])
        with torch.no_grad():
            fd = (loss_of(xp, stable) - loss_of(xm, stable)).item() / float(xp[b, k] - xm[b, k])
        xg = x.clone().requires_grad_(True)
        loss_of(xg, stable).backward()
        g = xg.grad[b, k].item()
        scale = max(abs(fd), 1e-3 * float(loss_of(x, stable)))
        err = abs(fd - g) / scale
        worst = max(worst, err)
        print(f"param ({b},{k}) fd {fd:.6e} autograd {g:.6e} rel {err:.2e}", flush=True)
        assert err < 2e-2, (b, k, fd, g)
    print("worst FD relative error", worst, flush=True)


def test_vertex_gradients_match_oracle(device):
    cfg = small_cfg()
    model = fit.ShapeModel(cfg, device)
    p = fit.sample_params(np.random.Generator(np.random.PCG64(5)), cfg.batch, cfg.n_basis, device)
    with torch.no_grad():
        vpos = model.vertices(p).contiguous()
    out = model.rast.forward(vpos, model.colors.expand(cfg.batch, model.N, 3).contiguous(),
                             model._tex.expand(cfg.batch, 1, 1, 3).contiguous(),
                             p.sh[:, None, :].expand(cfg.batch, cfg.views, 27).contiguous(),
                             model.extr.expand(cfg.batch, cfg.views, 12).reshape(cfg.batch, -1),
                             model.intr.expand(cfg.batch, cfg.views, 9).reshape(cfg.batch, -1))
    g = torch.randn_like(out[2])
    gp, gc, gt, gs = model.rast.backward(g, None, None, vpos, model.colors.expand(cfg.batch, model.N, 3).contiguous(),
                                         model._tex.expand(cfg.batch, 1, 1, 3).contiguous(),
                                         p.sh[:, None, :].expand(cfg.batch, cfg.views, 27).contiguous(), out[3], out[1],
                                         model.extr.expand(cfg.batch, cfg.views, 12).reshape(cfg.batch, -1),
                                         model.intr.expand(cfg.batch, cfg.views, 9).reshape(cfg.batch, -1))
    orc = O.Oracle(model.faces.reshape(-1), model.uv.reshape(-1), model.N, cfg.views, cfg.resolution, cfg.resolution,
                   albedo_mode="vertexColor", shading_mode="shaded")
    c = lambda t: t.detach().cpu().numpy()
    vcol = np.broadcast_to(c(model.colors), (cfg.batch, model.N, 3))
    sh = np.broadcast_to(c(p.sh)[:, None, :], (cfg.batch, cfg.views, 27))
    extr = np.broadcast_to(c(model.extr)[None], (cfg.batch, cfg.views, 12))
    intr = np.broadcast_to(c(model.intr)[None], (cfg.batch, cfg.views, 9))
    tex = np.full((cfg.batch, 1, 1, 3), 0.5, np.float32)
    ref = orc.forward(c(vpos), vcol, tex, sh, extr, intr)
    assert np.array_equal(ref["face"], c(out[1]))
    gr = orc.backward(c(g), np.zeros(ref["render"].shape, np.float32), np.zeros(ref["render"].shape, np.float32), c(vpos),
                      vcol, tex, sh, ref["vnormal"], ref["bary"], ref["face"], extr, intr)
    for name, x in (("pos", gp), ("col", gc), ("sh", gs)):
        e = rel_err(c(x), gr[name])
        print(f"fit scene grad {name} rel err {e:.3e}", flush=True)
        assert e <= 1e-4


def test_short_plain_fit_reduces_loss(device):
    cfg = small_cfg()
    model = fit.ShapeModel(cfg, device)
    gt = fit.sample_params(np.random.Generator(np.random.PCG64(6)), cfg.batch, cfg.n_basis, device)
    target, mask = fit.make_targets(model, gt)
    res = fit.fit(model, target, mask, fit.plain_init(cfg.batch, cfg.n_basis, device), iters=60)
    print("plain fit loss", res.losses[0], "->", res.losses[-1], flush=True)
    assert res.losses[-1] < 0.7 * res.losses[0]


@pytest.mark.parametrize("B,shape", [(1, (1, 1, 1)), (3, (2, 37, 53)), (16, (4, 64, 64)), (2, (1, 300, 301))])
def test_masked_squared_error_kernel(device, B, shape):
    """The fit's loss kernel (gvv_masked_sq_error and its gradient) against the same expression
    in float64 torch: per-subject sums within 1e-5 relative, the render gradient within 1e-6
    of the largest entry; ragged pixel counts (a partial last workgroup) and one pixel."""
    g = torch.Generator(device="cpu").manual_seed(sum(shape) + B)
    r = torch.rand((B,) + shape + (3,), generator=g).to(device)
    t = torch.rand((B,) + shape + (3,), generator=g).to(device)
    m = (torch.rand((B,) + shape + (1,), generator=g) > 0.3).float().to(device)
    m[0].view(-1)[0] = 1.0
    rr = r.clone().requires_grad_(True)
    sq = fit.masked_sq_error(rr, t, m)
    w = torch.rand(B, generator=g).to(device)
    (sq * w).sum().backward()
    d = (r.double() - t.double()) * m.double()
    ref = (d * d).flatten(1).sum(1)
    err = ((sq.double() - ref).abs() / ref.abs().clamp_min(1e-30)).max().item()
    gref = 2.0 * d * m.double() * w.double().view((B,) + (1,) * (r.dim() - 1))
    gerr = ((rr.grad.double() - gref).abs().max() / gref.abs().max().clamp_min(1e-30)).item()
    print(f"masked sq error B={B} {shape}: sum rel err {err:.2e}, grad rel err {gerr:.2e}", flush=True)
    assert err <= 1e-5 and gerr <= 1e-6


def test_masked_squared_error_is_deterministic(device):
    """Fixed-order sums: two evaluations of the same inputs agree bit for bit."""
    g = torch.Generator(device="cpu").manual_seed(11)
    r = torch.rand((4, 3, 128, 128, 3), generator=g).to(device)
    t = torch.rand((4, 3, 128, 128, 3), generator=g).to(device)
    m = (torch.rand((4, 3, 128, 128, 1), generator=g) > 0.5).float().to(device)
    a = fit.masked_sq_error(r, t, m)
    b = fit.masked_sq_error(r, t, m)
    assert torch.equal(a, b)


def test_masked_squared_error_rejects_mismatched_shapes(device):
    r = torch.zeros((2, 4, 4, 3), device=device)
    with pytest.raises(ValueError):
        fit.masked_sq_error(r, torch.zeros((2, 4, 5, 3), device=device), torch.zeros((2, 4, 4, 1), device=device))
    with pytest.raises(ValueError):
        fit.masked_sq_error(r, r, torch.zeros((2, 4, 4, 3), device=device))


@pytest.mark.parametrize("B", [1, 3, 17])
def test_shape_kernels_match_torch(device, B):
    """Posed vertices and the gradients of shape coefficients and pose from the shape-model
    kernels (gvv_shape_forward / gvv_shape_backward) against the same model in float64 torch
    operations: 3,002 vertices (a partial last chunk of 64), 8 basis vectors, B subjects
    including more than one 16-subject group; subject 0 has no rotation and subject 1 (when
    there is one) a rotation of 1e-5 rad, both on the series branch of Rodrigues' formula."""
    cfg = small_cfg()
    model = fit.ShapeModel(cfg, device)
    p = fit.sample_params(np.random.Generator(np.random.PCG64(20 + B)), B, cfg.n_basis, device)
    p.pose[0, :3] = 0.0
    if B > 1:
        p.pose[1, :3] = torch.tensor([1e-5, -2e-6, 3e-6], device=device)
    x = p.flat().detach().clone().requires_grad_(True)
    v = model.vertices(fit.Params.from_flat(x, cfg.n_basis))
    g = torch.randn(v.shape, generator=torch.Generator(device="cpu").manual_seed(B)).to(device)
    (v * g).sum().backward()

    xd = p.flat().detach().double().clone().requires_grad_(True)
    pd = fit.Params.from_flat(xd, cfg.n_basis)
    X = model.template.double() + (pd.alpha @ model.basis.double()).view(B, model.N, 3)
    m = model.motion(pd)
    R, t = m[:, :9].view(B, 3, 3), m[:, 9:]
    vd = torch.matmul(X - model.centre.double(), R.transpose(1, 2)) + model.centre.double() + t[:, None, :]
    (vd * g.double()).sum().backward()
    verr = ((v.double() - vd).abs().max() / vd.abs().max()).item()
    gerr = ((x.grad.double() - xd.grad).abs().max(1).values / xd.grad.abs().max(1).values).max().item()
    print(f"shape kernels B={B}: vertices rel err {verr:.2e}, parameter grad rel err {gerr:.2e}", flush=True)
    assert verr <= 1e-6 and gerr <= 1e-4
    with torch.no_grad():
        assert (model.vertices(p) - model.vertices_reference(p)).abs().max() <= 1e-5 * vd.abs().max()
