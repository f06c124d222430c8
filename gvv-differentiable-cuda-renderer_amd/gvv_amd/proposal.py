"""Proposal network: predicts shape, SH and pose coefficients from renders to warm-start the fit.

Architecture (all convolutions 3x3, stride 2, pad 1, ReLU; NHWC bf16 activations):

    input  256 x 256 RGB   the 4 views as one 256 x 256 render: each box-filtered from the fit
                           resolution to 128 x 128 and tiled 2 x 2, RGB - 0.5, stored as 8 bf16
                           channels of which 5 are zero (features=0, the default); features=1
                           adds the covered fraction and tile-local x and y
    conv1  128 x 128 x 32
    conv2   64 x  64 x 64
    conv3   32 x  32 x 128
    conv4   16 x  16 x 128
    conv5    8 x   8 x 256
    conv6    4 x   4 x 256
    head   linear 4096 -> 113 (80 shape, 27 SH, 6 pose), run as a 1 x 1 convolution over a
           1 x 1 image with 4096 channels, padded to 128 outputs (the pad rows stay zero)

Forward and backward run entirely on the hand-written BF16 MFMA kernels of libgvv_amd.so
(include/gvv_proposal.h): FP32 master weights and Adam moments, bf16 weight copies refreshed
after each step, FP32 accumulation.  Training data are renders the rasterizer generates from
parameters drawn from the fit workload's priors (gvv_amd.fit.sample_params).

The reference has no learned component; its fit loop (python/test_gradients_VertexColor.py:93-138)
is what the proposal warm-starts.
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, fit

N_OUT = 113
N_OUT_PAD = 128
LAYERS = ((8, 32), (32, 64), (64, 128), (128, 128), (128, 256), (256, 256))   # (Cin, Cout), 3x3 s2 p1
INPUT_RES = 256
VIEW_RES = 128
FEATURES = 0                      # input channels: 0 = RGB only; 1 = RGB - 0.5, coverage, tile x/y
TRAIN_EPOCHS = 4                  # passes over the stored training renders


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def param_stats(n_basis: int, device):
    """Mean and spread of each output coefficient under the fit workload's priors
    (gvv_amd.fit.sample_params): the network predicts (param - mean) / std."""
    mean = torch.zeros(n_basis + 33, device=device)
    std = torch.ones(n_basis + 33, device=device)
    std[:n_basis] = 0.5
    sh_mean = torch.zeros(27, device=device)
    sh_std = torch.full((27,), 0.1, device=device)
    for c in (0, 9, 18):
        sh_mean[c] = 0.8
        sh_std[c] = 0.4 / math.sqrt(12.0)
    mean[n_basis:n_basis + 27] = sh_mean
    std[n_basis:n_basis + 27] = sh_std
    std[n_basis + 27:n_basis + 30] = 0.1
    std[n_basis + 30:n_basis + 33] = 0.3
    return mean, std


@dataclass
class _Layer:
    H: int            # input spatial size (square)
    Cin: int
    Cout: int
    ks: int
    stride: int
    pad: int
    relu: bool

    @property
    def OH(self):
        return (self.H + 2 * self.pad - self.ks) // self.stride + 1


class ProposalNet:
    """Conv encoder + linear head on the BF16 MFMA kernels.  `forward` caches the activations
    that `backward` needs; `backward` writes FP32 gradients of the FP32 master weights
    (self.gw, self.gb); `refresh` re-derives the bf16 copies after an optimizer step."""

    def __init__(self, device=None, seed: int = 0, features: int = FEATURES):
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.features = int(features)
        self.in_used = 6 if self.features else 3          # live input channels of the 8
        self.L = _lib.load()
        layers, H = [], INPUT_RES
        for cin, cout in LAYERS:
            layers.append(_Layer(H, cin, cout, 3, 2, 1, True))
            H = layers[-1].OH
        self.flat = H * H * LAYERS[-1][1]
        layers.append(_Layer(1, self.flat, N_OUT_PAD, 1, 1, 0, False))
        self.layers = layers
        g = torch.Generator(device="cpu").manual_seed(seed)
        self.w, self.b = [], []
        for i, l in enumerate(layers):
            if i == len(layers) - 1:
                std = math.sqrt(1.0 / l.Cin)
            else:
                fan_in = l.ks * l.ks * (self.in_used if i == 0 else l.Cin)
                std = math.sqrt(2.0 / fan_in)
            w = torch.randn((l.Cout, l.ks, l.ks, l.Cin), generator=g) * std
            if i == 0:
                w[..., self.in_used:] = 0.0       # padding input channels
            if i == len(layers) - 1:
                w[N_OUT:] = 0.0                   # padded outputs
            self.w.append(w.to(self.device).contiguous())
            self.b.append(torch.zeros(l.Cout, device=self.device))
        self.gw = [torch.zeros_like(w) for w in self.w]
        self.gb = [torch.zeros_like(b) for b in self.b]
        self.wb = [torch.empty(w.shape, dtype=torch.bfloat16, device=self.device) for w in self.w]
        self.wt = [torch.empty((l.Cin, l.ks, l.ks, l.Cout), dtype=torch.bfloat16, device=self.device) for l in layers]
        self._ws = torch.empty(0, device=self.device)
        self._acts = None
        self.refresh()

    # -- parameters -------------------------------------------------------------------------
    def parameters(self):
        return self.w + self.b

    def refresh(self):
        """Re-derive the bf16 (and transposed bf16) weight copies from the FP32 masters."""
        s = _stream()
        for l, w, wb, wt in zip(self.layers, self.w, self.wb, self.wt):
            _lib.check(self.L.gvv_conv_prepare_weights(l.Cout, l.ks * l.ks, l.Cin, _ptr(w), _ptr(wb), _ptr(wt), s),
                       "gvv_conv_prepare_weights")

    def state_dict(self):
        return {f"w{i}": w.detach().cpu() for i, w in enumerate(self.w)} | \
               {f"b{i}": b.detach().cpu() for i, b in enumerate(self.b)} | \
               {"features": torch.tensor(self.features)}

    def load_state_dict(self, sd):
        if "features" in sd and int(sd["features"]) != self.features:
            raise ValueError(f"weights were trained with features={int(sd['features'])}, net has {self.features}")
        for i in range(len(self.layers)):
            self.w[i].copy_(sd[f"w{i}"].to(self.device, torch.float32))
            self.b[i].copy_(sd[f"b{i}"].to(self.device, torch.float32))
        self.refresh()

    # -- forward / backward -----------------------------------------------------------------
    def forward(self, x):
        """x: [B, 256, 256, 8] bf16 NHWC.  Returns [B, 113] float32 (normalised coefficients)."""
        B = x.shape[0]
        assert x.dtype == torch.bfloat16 and x.is_contiguous() and tuple(x.shape[1:]) == (INPUT_RES, INPUT_RES, 8)
        s = _stream()
        acts = [x]
        for i, l in enumerate(self.layers):
            last = i == len(self.layers) - 1
            if last:
                y = torch.empty((B, N_OUT_PAD), dtype=torch.float32, device=self.device)
            else:
                y = torch.empty((B, l.OH, l.OH, l.Cout), dtype=torch.bfloat16, device=self.device)
            _lib.check(self.L.gvv_conv_forward(B, l.H, l.H, l.Cin, l.Cout, l.ks, l.stride, l.pad, _ptr(acts[-1]),
                                               _ptr(self.wb[i]), _ptr(self.b[i]), int(l.relu), _ptr(y), int(last), s),
                       "gvv_conv_forward")
            acts.append(y)
        self._acts = acts
        return acts[-1][:, :N_OUT]

    def backward(self, dout):
        """dout: [B, 113] float32 gradient of the loss at the output.  Fills self.gw / self.gb."""
        acts = self._acts
        assert acts is not None, "backward before forward"
        B = dout.shape[0]
        s = _stream()
        dz = torch.zeros((B, N_OUT_PAD), dtype=torch.bfloat16, device=self.device)
        dz[:, :N_OUT] = dout.to(torch.bfloat16)
        for i in range(len(self.layers) - 1, -1, -1):
            l = self.layers[i]
            x = acts[i]
            need = self.L.gvv_conv_workspace_floats(B, l.H, l.H, l.Cin, l.Cout, l.ks, l.stride, l.pad)
            if need < 0:
                raise _lib.GvvError("invalid layer geometry")
            if self._ws.numel() < need:
                self._ws = torch.empty(int(need), device=self.device)
            _lib.check(self.L.gvv_conv_backward_weight(B, l.H, l.H, l.Cin, l.Cout, l.ks, l.stride, l.pad, _ptr(x), _ptr(dz),
                                                       _ptr(self.gw[i]), _ptr(self.gb[i]), _ptr(self._ws),
                                                       self._ws.numel(), s), "gvv_conv_backward_weight")
            if i == 0:
                break
            dx = torch.empty(x.shape, dtype=torch.bfloat16, device=self.device)
            _lib.check(self.L.gvv_conv_backward_data(B, l.H, l.H, l.Cin, l.Cout, l.ks, l.stride, l.pad, _ptr(dz),
                                                     _ptr(self.wt[i]), _ptr(x), _ptr(dx), s), "gvv_conv_backward_data")
            dz = dx
        self.gw[0][..., self.in_used:] = 0.0
        self.gw[-1][N_OUT:] = 0.0
        self.gb[-1][N_OUT:] = 0.0


def pack_mosaic(render, cols: int | None = None, features: int = FEATURES):
    """[B, V, R, R, 3] float32 renders -> the network input [B, 256, 256, 8] bf16: views box-filtered
    to 128 x 128 and tiled (view v at tile (v // cols, v % cols))."""
    B, V, R = render.shape[0], render.shape[1], render.shape[2]
    cols = cols or int(math.ceil(math.sqrt(V)))
    rows = (V + cols - 1) // cols
    if R % VIEW_RES or cols * VIEW_RES > INPUT_RES or rows * VIEW_RES > INPUT_RES:
        raise ValueError(f"{V} views of {R} x {R} do not pack into a {INPUT_RES} x {INPUT_RES} mosaic of "
                         f"{VIEW_RES} x {VIEW_RES} tiles")
    render = render.contiguous()
    x = torch.empty((B, INPUT_RES, INPUT_RES, 8), dtype=torch.bfloat16, device=render.device)
    if cols * VIEW_RES < INPUT_RES or rows * VIEW_RES < INPUT_RES:
        x.zero_()
        view = x[:, :rows * VIEW_RES, :cols * VIEW_RES]
        tmp = torch.empty((B, rows * VIEW_RES, cols * VIEW_RES, 8), dtype=torch.bfloat16, device=render.device)
        _lib.check(_lib.load().gvv_pack_mosaic(_ptr(render), B, V, R, R, cols, R // VIEW_RES, features, _ptr(tmp),
                                               _stream()),
                   "gvv_pack_mosaic")
        view.copy_(tmp)
        return x
    _lib.check(_lib.load().gvv_pack_mosaic(_ptr(render), B, V, R, R, cols, R // VIEW_RES, features, _ptr(x), _stream()),
               "gvv_pack_mosaic")
    return x


def render_inputs(model: fit.ShapeModel, p: fit.Params, features: int = FEATURES):
    """Network input for parameters p: the fit workload's own render (all views at the fit
    resolution), packed by pack_mosaic -- the same path the fit's target images take."""
    with torch.no_grad():
        render = model.render(p)[0]
    return pack_mosaic(render, features=features)


@dataclass
class TrainLog:
    samples: int
    seconds: float
    losses: list


def train(net: ProposalNet, model: fit.ShapeModel, samples: int = 200_000, batch: int = 16, lr: float = 7e-4,
          seed: int = 1, log_every: int = 0, warmup: float = 0.02, weights=(1.0, 1.0, 1.0),
          epochs: int = 1) -> TrainLog:
    """Trains on `samples` renders generated from the fit priors: Adam, linear warm-up over the
    first `warmup` fraction of the steps then cosine decay, weighted MSE on normalised
    coefficients (`weights` for the shape, SH and pose groups).

    epochs = 1: each render is generated on the fly and used once.  epochs > 1: the renders'
    RGB channels (bf16, 384 KiB per render at the fit's 4 views) and normalised coefficients
    are generated once and kept in device memory, and every epoch visits them all in a seeded
    random order of batches (RGB input only, features = 0)."""
    dev = model.device
    K = model.cfg.n_basis
    mean, std = param_stats(K, dev)
    w = torch.ones(K + 33, device=dev)
    w[:K], w[K:K + 27], w[K + 27:] = weights[0], weights[1], weights[2]
    w = w * ((K + 33) / w.sum())
    opt = torch.optim.Adam(net.parameters(), lr=lr)
    nb = max(1, samples // batch)
    if epochs < 1:
        raise ValueError("epochs must be >= 1")
    if epochs > 1 and net.features != 0:
        raise ValueError("training over stored renders keeps RGB only: use features=0")
    steps = nb * epochs
    n_warm = max(1, int(warmup * steps))

    def factor(step):
        if step < n_warm:
            return (step + 1) / n_warm
        return 0.5 * (1.0 + math.cos(math.pi * min(step - n_warm, steps - n_warm) / max(1, steps - n_warm)))

    sched = torch.optim.lr_scheduler.LambdaLR(opt, factor)
    rng = np.random.Generator(np.random.PCG64(seed))
    losses = []
    t0 = time.perf_counter()

    def sgd_step(step, x, target):
        out = net.forward(x)
        diff = out - target
        loss = (diff * diff).mean()
        net.backward(diff * w * (2.0 / diff.numel()))
        for prm, g in zip(net.parameters(), net.gw + net.gb):
            prm.grad = g
        opt.step()
        sched.step()
        net.refresh()
        if log_every and (step % log_every == 0 or step == steps - 1):
            losses.append((step, float(loss)))
            print(f"proposal step {step}/{steps} loss {float(loss):.4f} {time.perf_counter() - t0:.1f}s", flush=True)

    if epochs == 1:
        for step in range(steps):
            p = fit.sample_params(rng, batch, K, dev)
            sgd_step(step, render_inputs(model, p, net.features), (p.flat() - mean) / std)
    else:
        X = torch.empty((nb, batch, INPUT_RES, INPUT_RES, 3), dtype=torch.bfloat16, device=dev)
        Y = torch.empty((nb, batch, K + 33), dtype=torch.float32, device=dev)
        for i in range(nb):
            p = fit.sample_params(rng, batch, K, dev)
            X[i] = render_inputs(model, p, 0)[..., :3]
            Y[i] = (p.flat() - mean) / std
        x = torch.zeros((batch, INPUT_RES, INPUT_RES, 8), dtype=torch.bfloat16, device=dev)
        order = np.random.Generator(np.random.PCG64(seed + 1))
        step = 0
        for _ in range(epochs):
            for i in order.permutation(nb):
                x[..., :3] = X[int(i)]
                sgd_step(step, x, Y[int(i)])
                step += 1
        del X, Y
    torch.cuda.synchronize()
    return TrainLog(samples=nb * batch, seconds=time.perf_counter() - t0, losses=losses)


def propose(net: ProposalNet, model: fit.ShapeModel, target_render) -> fit.Params:
    """Parameters predicted from target images [B, V, R, R, 3]."""
    mean, std = param_stats(model.cfg.n_basis, model.device)
    out = net.forward(pack_mosaic(target_render, features=net.features))
    return fit.Params.from_flat(mean + std * out, model.cfg.n_basis)


def subject_losses(model: fit.ShapeModel, p: fit.Params, target, mask):
    """Per-subject masked photometric loss [B] (the fit's objective split by subject)."""
    with torch.no_grad():
        r = model.render(p)[0]
    d = (r - target) * mask
    return (d * d).flatten(1).sum(1) / mask.flatten(1).sum(1).clamp_min(1.0)


def warm_start(net: ProposalNet, model: fit.ShapeModel, target, mask):
    """Proposal for each subject, replaced by the plain initialisation wherever the proposal's
    own loss is worse.  Returns (params, number of subjects that fell back)."""
    B = target.shape[0]
    prop = propose(net, model, target)
    plain = fit.plain_init(B, model.cfg.n_basis, model.device)
    worse = subject_losses(model, prop, target, mask) > subject_losses(model, plain, target, mask)
    x = torch.where(worse[:, None], plain.flat(), prop.flat())
    return fit.Params.from_flat(x, model.cfg.n_basis), int(worse.sum())


@dataclass
class WarmStartReport:
    plain_final_loss: float
    plain_seconds: float                # the plain fit's wall-clock for all its iterations
    plain_seconds_to_final: float       # until the plain loss first reached its final value
    plain_iters: int
    warm_initial_loss: float
    plain_initial_loss: float
    warm_seconds: float | None          # until the warm fit reached plain_final_loss (incl. proposal)
    warm_iters: int | None
    fallbacks: int
    warm_runs: list | None = None       # (seconds, iterations) of every warm run; warm_seconds is the median
    plain_runs: list | None = None      # (seconds, final loss) of every plain run

    @property
    def ratio(self):
        return None if self.warm_seconds is None else self.warm_seconds / self.plain_seconds

    @property
    def ratios(self):
        return [None if w is None else w / self.plain_seconds for w, _ in (self.warm_runs or [])]


def warm_fit(net: ProposalNet, model: fit.ShapeModel, target, mask, iters: int, level: float):
    """The warm-started fit, timed from the proposal on: the network's prediction from the
    target images, the plain start's per-subject losses (one single-subject render), then the
    fit, which takes each subject's loss at the proposal from its own first forward and falls
    back to the plain start where that is worse."""
    t0 = time.perf_counter()
    prop = propose(net, model, target)
    B = target.shape[0]
    plain = fit.plain_init(B, model.cfg.n_basis, model.device)
    res = fit.fit(model, target, mask, prop, iters=iters, stop_at=level, t0=t0,
                  fallback=(plain, fit.plain_subject_losses(model, target, mask)))
    torch.cuda.synchronize()
    it = next((i for i, l in enumerate(res.losses) if l <= level), None)
    return res, res.time_to(level), it


def compare_fits(net: ProposalNet, model: fit.ShapeModel, gt: fit.Params, iters: int | None = None,
                 plain_runs: int = 5, warm_runs: int = 5) -> WarmStartReport:
    """Plain fit for `iters` Adam steps (`plain_runs` times: float atomics make its last bits, and
    so its final loss, vary from run to run; the run with the median final loss is the
    reference), then the warm-started fit (`warm_runs` times) until it reaches that run's final
    loss (at most the same number of steps).  Both clocks include everything done on the way
    (the warm one includes the proposal, its fallback check and packing); the reported warm time
    is the median run's."""
    target, mask = fit.make_targets(model, gt)
    B = target.shape[0]
    iters = model.cfg.iters if iters is None else iters
    fit.fit(model, target, mask, fit.plain_init(B, model.cfg.n_basis, model.device), iters=2)   # warm caches
    torch.cuda.synchronize()
    runs = []
    for _ in range(max(1, plain_runs)):
        runs.append(fit.fit(model, target, mask, fit.plain_init(B, model.cfg.n_basis, model.device), iters=iters))
        torch.cuda.synchronize()
    order = sorted(range(len(runs)), key=lambda i: runs[i].final_loss)
    plain = runs[order[len(runs) // 2]]
    level = plain.final_loss
    warm = []
    for _ in range(max(1, warm_runs)):
        warm.append(warm_fit(net, model, target, mask, iters, level))
    timed = sorted(warm, key=lambda w: float("inf") if w[1] is None else w[1])
    res, t_warm, it_warm = timed[len(timed) // 2]
    return WarmStartReport(plain_final_loss=level, plain_seconds=plain.times[-1],
                           plain_seconds_to_final=plain.time_to(level), plain_iters=iters,
                           warm_initial_loss=res.losses[0], plain_initial_loss=plain.losses[0],
                           warm_seconds=t_warm, warm_iters=it_warm, fallbacks=res.fallbacks,
                           warm_runs=[(w[1], w[2]) for w in warm],
                           plain_runs=[(r.times[-1], r.final_loss) for r in runs])
