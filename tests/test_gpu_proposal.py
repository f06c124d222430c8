"""The proposal network on the BF16 MFMA kernels: whole-network gradients against FP64 CPU
autograd of the same network (same bf16 weights, bf16 activations), training on rasterizer
renders, and the warm start against the plain initialisation."""
import math

import numpy as np
import pytest
import torch

from gvv_amd import fit, proposal
from oracle import conv_oracle as CO

pytestmark = pytest.mark.gpu


def bf16r(t):
    return t.to(torch.bfloat16).to(t.dtype)


def cpu_reference(net, x_bf16, dout):
    """FP64 CPU autograd of the network with the GPU path's roundings: bf16 weights, bf16
    activations between layers, bf16 gradient at the head."""
    ws = [w.to(torch.bfloat16).double().cpu().requires_grad_(True) for w in net.w]
    bs = [b.double().cpu().requires_grad_(True) for b in net.b]
    h = x_bf16.double().cpu().permute(0, 3, 1, 2)
    for i, l in enumerate(net.layers[:-1]):
        z = torch.nn.functional.conv2d(h, ws[i].permute(0, 3, 1, 2), bs[i], stride=l.stride, padding=l.pad)
        h = torch.relu(z)
        h = h + (bf16r(h.detach().float()).double() - h.detach())     # bf16 storage, straight-through
    flat = h.permute(0, 2, 3, 1).reshape(h.shape[0], -1)
    out = flat @ ws[-1].reshape(ws[-1].shape[0], -1).t() + bs[-1]
    g = torch.zeros_like(out)
    g[:, :proposal.N_OUT] = dout.cpu().double().to(torch.bfloat16).double()
    out.backward(g)
    return out[:, :proposal.N_OUT].detach(), [w.grad for w in ws], [b.grad for b in bs]


def test_network_gradients_vs_fp64(device):
    net = proposal.ProposalNet(device, seed=3)
    rng = np.random.default_rng(0)
    x = torch.from_numpy(CO.bf16_round(rng.random((2, 256, 256, 8)).astype(np.float32))).to(device)
    x[..., 3:] = 0
    x = x.to(torch.bfloat16).contiguous()
    out = net.forward(x)
    dout = torch.from_numpy(rng.standard_normal((2, proposal.N_OUT)).astype(np.float32)).to(device)
    net.backward(dout)
    ref_out, ref_gw, ref_gb = cpu_reference(net, x, dout)
    e = float((out.double().cpu() - ref_out).norm() / ref_out.norm())
    print(f"output rel err {e:.2e}", flush=True)
    assert e < 2e-2
    for i in range(len(net.layers)):
        gw = net.gw[i].double().cpu()
        gb = net.gb[i].double().cpu()
        rw, rb = ref_gw[i], ref_gb[i]
        if i == 0:
            rw = rw.clone()
            rw[..., 3:] = 0
        if i == len(net.layers) - 1:
            rw, rb = rw.clone(), rb.clone()
            rw[proposal.N_OUT:] = 0
            rb[proposal.N_OUT:] = 0
        ew = float((gw - rw).norm() / rw.norm())
        eb = float((gb - rb).norm() / max(float(rb.norm()), 1e-30))
        print(f"layer {i}: dW rel err {ew:.2e}  db rel err {eb:.2e}", flush=True)
        assert ew < 5e-2 and eb < 5e-2


def small_model(device):
    cfg = fit.FitConfig(cols=60, rows=50, views=4, resolution=256, batch=8, n_basis=80)
    return fit.ShapeModel(cfg, device)


@pytest.fixture(scope="module")
def trained(device):
    """A short training run on rasterizer renders of the reduced fit workload."""
    model = small_model(device)
    net = proposal.ProposalNet(device, seed=0)
    log = proposal.train(net, model, samples=32 * 300, batch=32, lr=1e-3, seed=5, log_every=1)
    return model, net, log


def test_training_beats_the_prior_mean(trained):
    """The loss is the MSE of normalised coefficients, so predicting the prior mean scores 1.
    A short training run on rasterizer renders must do clearly better than that."""
    model, net, log = trained
    first = np.mean([l for _, l in log.losses[:5]])
    last = np.mean([l for _, l in log.losses[-20:]])
    print(f"training loss {first:.4f} -> {last:.4f} in {log.seconds:.1f}s", flush=True)
    assert math.isfinite(last) and last < 0.9


def test_training_over_stored_renders(device):
    """epochs > 1: the renders are generated once, kept in device memory and visited every
    epoch; the run takes epochs x batches steps and learns as the on-the-fly run does."""
    model = small_model(device)
    net = proposal.ProposalNet(device, seed=0)
    log = proposal.train(net, model, samples=32 * 100, batch=32, lr=1e-3, seed=5, log_every=1, epochs=3)
    steps = [s for s, _ in log.losses]
    assert steps[-1] == 299 and log.samples == 32 * 100
    first = np.mean([l for _, l in log.losses[:5]])
    last = np.mean([l for _, l in log.losses[-20:]])
    print(f"stored renders, 3 epochs: loss {first:.4f} -> {last:.4f} in {log.seconds:.1f}s", flush=True)
    assert math.isfinite(last) and last < 0.9
    with pytest.raises(ValueError):
        proposal.train(proposal.ProposalNet(device, seed=0, features=1), model, samples=64, batch=32, epochs=2)


def test_trained_network_gains_over_the_plain_fit(device, trained):
    """What the trained network gains over the method without it, on held-out subjects (a
    seed the training never drew): the warm start begins below the plain start's loss, and the
    warm-started fit reaches the plain fit's final loss (after 100 Adam steps) in far fewer
    steps than the plain fit takes."""
    model, net, _ = trained
    gt = fit.sample_params(np.random.Generator(np.random.PCG64(424242)), model.cfg.batch, model.cfg.n_basis, device)
    rep = proposal.compare_fits(net, model, gt, iters=100, plain_runs=1, warm_runs=1)
    print(f"plain: loss {rep.plain_initial_loss:.4f} -> {rep.plain_final_loss:.4f} in {rep.plain_iters} steps, "
          f"{rep.plain_seconds:.3f}s; warm: starts at {rep.warm_initial_loss:.4f}, reaches it after "
          f"{rep.warm_iters} steps, {rep.warm_seconds}s (ratio {rep.ratio}), fallbacks {rep.fallbacks}", flush=True)
    assert rep.warm_initial_loss < rep.plain_initial_loss
    assert rep.warm_iters is not None and rep.warm_iters <= 0.6 * rep.plain_iters
    assert rep.ratio is not None and rep.ratio < 0.8


def test_warm_start_fallback_never_worse(device):
    """With an untrained network the fallback must keep every subject at least as good as the
    plain initialisation."""
    model = small_model(device)
    net = proposal.ProposalNet(device, seed=1)
    gt = fit.sample_params(np.random.Generator(np.random.PCG64(11)), 4, model.cfg.n_basis, device)
    target, mask = fit.make_targets(model, gt)
    init, fallbacks = proposal.warm_start(net, model, target, mask)
    plain = fit.plain_init(4, model.cfg.n_basis, device)
    li = proposal.subject_losses(model, init, target, mask)
    lp = proposal.subject_losses(model, plain, target, mask)
    print("fallbacks", fallbacks, li.tolist(), lp.tolist(), flush=True)
    assert torch.all(li <= lp)
    # the fit's own fallback (decided from its first forward) starts no worse than plain either
    res, _, _ = proposal.warm_fit(net, model, target, mask, iters=3, level=-1.0)
    plain_loss = float(fit.masked_l2(model.render(plain)[0].detach(), target, mask))
    print("fit fallbacks", res.fallbacks, "first loss", res.losses[0], "plain", plain_loss, flush=True)
    assert res.fallbacks == fallbacks
    assert res.losses[0] <= plain_loss * (1 + 1e-5)
