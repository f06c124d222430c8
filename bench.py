#!/usr/bin/env python3
"""Throughput of the differentiable rasterizer, forward + backward, on one MI355X.

Workload: the capped cylinder of 100,000 triangles / 50,002 vertices, 1024^2 smooth-noise
texture, 8 views at 1024x1024 on a 3.5 m ring, textured + shaded, full backward to vertices,
texture and SH.  One step = one forward and one backward of the renderer over the 8 views with
every input already resident in HBM; the upstream gradient is the sum-of-squares loss gradient
2(render - target) of the first forward, fixed across steps.

Prints one JSON line: Mpix/s = 8 x 1024 x 1024 / (t_fwd + t_bwd) over K uninstrumented steps
(launch settings chosen first by timing the autotuner's candidates), plus the roofline of the
dominant kernel (HIP-event timing through the library's profiling hook over a second run of
the same K steps), the step's algorithmic bytes against HBM peak, the atomic rate against a
measured atomic peak, the CPU oracle timed
on a bounded sample of the same workload, and (unless --no-fit) the fit time-to-tolerance of
the proposal-warm-started fit against the plain fit:

  fit      35,002-vertex / 70,000-triangle template, 80-dim shape basis, 27 SH, 6-DoF pose;
           16 subjects x 4 views at 512^2, masked L2, Adam lr 1e-2, 200 iterations (plain,
           zero-initialised); the proposal network is trained first on renders the renderer
           generates once and keeps in device memory (--train-samples, default 200,000; four
           passes over them), then the warm-started fit runs until it reaches the plain fit's
           final loss.  Reported: both wall-clock times and their ratio.

--workload multi instead times the batched uneven workload (32 meshes, 2k-200k triangles,
1-8 views at 256/512/1024^2) through the multi-mesh runtime: autotuned, one hipGraph launch
per forward+backward step.  --workload fit times iterations of the plain fit alone.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "gvv-differentiable-cuda-renderer_amd")
for _p in (PKG, ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
# wave64 VALU instructions/s: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64 instruction per 4 cycles on
# a 16-lane SIMD (78.6 TFLOP/s of FP32 FMAs; the 157.3 TFLOP/s FP32 figure counts packed-FP32
# instructions, two FMAs per lane, which this library is built without)
VALU_PEAK_PER_S = 256 * 4 * 2.4e9 / 4


def kernel_bytes(name, scene, covered, texW, texH):
    """Algorithmic bytes of one launch of `name` for this workload: every distinct byte the
    kernel's algorithm must read or write once.  Per-unit figures:
      k_pixel_grad  16 B/px (render grad 12 + face id 4) + per vertex 12 (pos) + 12 C (normals)
                    + 24 (pos and normal-gradient accumulators) + per face 36 (ids 12 + UVs 24)
                    + 24 B/texel (texture read 12 + texture-gradient write 12) + 108 B/view (SH)
      k_resolve     8 B/px key read + 24 B/px outputs (bary 8, face 4, render 12) + 8 B per
                    covered px (key reset) + per vertex 12 + 12 C + per face 36 + 12 B/texel
      k_raster      per face-view 12 (ids) + 36 (world vertices) + 48 (projected vertices)
                    + 8 B per covered px (visibility key)"""
    B, C, N, F = scene.B, scene.C, scene.N, scene.F
    P = B * C * scene.width * scene.height
    T = B * texW * texH
    if name == "k_pixel_grad":
        return 16 * P + B * N * (12 + 12 * C + 24) + 36 * F + 24 * T + 108 * B * C
    if name == "k_resolve":
        return 32 * P + 8 * covered + B * N * (12 + 12 * C) + 36 * F + 12 * T
    if name == "k_raster":
        return B * C * F * (12 + 36 + 48) + 8 * covered
    return None


def atomic_peak(dev, reps=3):
    """Float-atomic throughput of this GPU (atomics/s) from the library's calibration kernel:
    scattered (hashed) and streaming (consecutive lanes, consecutive addresses) patterns over a
    12 MiB buffer, best of `reps`."""
    import ctypes
    import torch
    from gvv_amd import _lib
    L = _lib.load()
    n, updates = 3 << 20, 1 << 26
    buf = torch.zeros(n, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = {}
    for name, scattered in (("scattered", 1), ("streaming", 0)):
        best = None
        for _ in range(reps + 1):
            a, z = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            _lib.check(L.gvv_calibrate_atomics(ctypes.c_void_p(buf.data_ptr()), n, updates, scattered, stream),
                       "gvv_calibrate_atomics")
            z.record()
            torch.cuda.synchronize()
            ms = a.elapsed_time(z)
            best = ms if best is None else min(best, ms)
        out[name] = updates / (best / 1e3)
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene, threads, views=8, reps=3, g_render=None):
    """Oracle forward + backward over `views` of the views, multithreaded: one warm-up, then the
    median of `reps` timed repetitions (one-time setup excluded).  With all views and a given
    upstream gradient `g_render`, the backward takes that gradient (the one the GPU step used),
    so that the last repetition's outputs can be compared with the GPU's.  Returns the baseline
    dict and (forward outputs, gradients) of the last repetition."""
    import numpy as np
    from oracle import oracle as O
    O.set_threads(threads)
    kw = scene.renderer_kwargs()
    views = min(views, scene.C)
    kw["n_cameras"] = views
    orc = O.Oracle(**kw)
    ins = (scene.vpos, scene.vcol, scene.tex, np.ascontiguousarray(scene.sh[:, :views]),
           np.ascontiguousarray(scene.extr[:, :views]), np.ascontiguousarray(scene.intr[:, :views]))
    target = np.ascontiguousarray(scene.target[:, :views])
    times = []
    out = gout = None
    for rep in range(reps + 1):
        t0 = time.perf_counter()
        out = orc.forward(*ins)
        t1 = time.perf_counter()
        if g_render is not None and views == scene.C:
            g = g_render
        else:
            g = (2.0 * (out["render"] - target)).astype(np.float32)
        t2 = time.perf_counter()
        gout = orc.backward(g, np.zeros_like(g), target, scene.vpos, scene.vcol, scene.tex, ins[3], out["vnormal"],
                            out["bary"], out["face"], ins[4], ins[5])
        t3 = time.perf_counter()
        if rep > 0:
            times.append((t1 - t0, t3 - t2))
    times.sort(key=lambda t: t[0] + t[1])
    t_fwd, t_bwd = times[len(times) // 2]
    px = views * scene.width * scene.height * scene.B
    return dict(value=px / (t_fwd + t_bwd) / 1e6, unit="Mpix/s", cores=threads, kind="port",
                sample=f"FP64 oracle fwd+bwd on {views} of {scene.C} views ({px} px, {scene.F} triangles), median of "
                       f"{reps} after 1 warm-up: fwd {t_fwd:.2f}s bwd {t_bwd:.2f}s; {threads} threads on {cpu_model()}"), \
        (out, gout, views == scene.C)


# accuracy bar of the path: ids, barycentrics and vertex normals bit-exact; render within 1e-5
# and gradients within 1e-4 max-norm relative of the FP64 oracle
RENDER_TOL, GRAD_TOL, ELEM_FLOOR = 1e-5, 1e-4, 1e-3


def rel_errors(x, ref):
    """Max-norm relative error and, over the entries with |ref| >= 1e-3 max|ref|, the max,
    99.9th percentile and median of the elementwise relative error."""
    import numpy as np
    x = np.asarray(x, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    amax = float(np.abs(ref).max()) if ref.size else 0.0
    diff = float(np.abs(x - ref).max()) if ref.size else 0.0
    r = dict(maxnorm=diff / amax if amax > 0 else diff)
    if amax > 0:
        m = np.abs(ref) >= ELEM_FLOOR * amax
        e = np.abs(x[m] - ref[m]) / np.abs(ref[m])
        r.update(elem_max=float(e.max()), elem_p999=float(np.quantile(e, 0.999)), elem_median=float(np.median(e)))
    return {k: float(f"{v:.3g}") for k, v in r.items()}


def accuracy_report(gpu, ref, gref):
    """Per-output errors of the GPU step's outputs against the oracle's on the same inputs."""
    import numpy as np
    acc = dict(face_mismatch=int((gpu["face"] != ref["face"]).sum()),
               bary_bit_exact=bool(np.array_equal(gpu["bary"].view(np.uint32), ref["bary"].view(np.uint32))),
               vnormal_bit_exact=bool(np.array_equal(gpu["vnormal"].view(np.uint32), ref["vnormal"].view(np.uint32))),
               render=rel_errors(gpu["render"], ref["render"]))
    for k in ("pos", "col", "tex", "sh"):
        acc["g_" + k] = rel_errors(gpu["g_" + k], gref[k])
    acc["tolerances"] = dict(render=RENDER_TOL, grad=GRAD_TOL, elementwise_floor=ELEM_FLOOR)
    acc["meets_tolerances"] = bool(acc["face_mismatch"] == 0 and acc["bary_bit_exact"] and acc["vnormal_bit_exact"]
                                   and acc["render"]["maxnorm"] <= RENDER_TOL
                                   and all(acc["g_" + k]["maxnorm"] <= GRAD_TOL for k in ("pos", "col", "tex", "sh")))
    return acc


def load_counters(workload):
    """Counter figures of the kernels recorded with rocprofv3 for this workload (copy-calibrated
    HBM traffic, VALU instructions and atomic requests per launch), with the launch settings and
    kernel-source digest they were recorded at; None if absent."""
    f = os.path.join(ROOT, "profiles", f"counters_{workload}.json")
    if not os.path.exists(f):
        return None
    with open(f) as fh:
        return json.load(fh)


def source_digest():
    """Digest of the HIP sources of the library: counters recorded from other sources are stale."""
    import hashlib
    h = hashlib.sha256()
    csrc = os.path.join(PKG, "csrc")
    for name in sorted(os.listdir(csrc)):
        with open(os.path.join(csrc, name), "rb") as fh:
            h.update(name.encode() + b"\0" + fh.read())
    return h.hexdigest()[:16]


def step_bytes(scene, covered, texW, texH):
    """Algorithmic bytes of one forward + backward (every distinct input byte read once, every
    output byte written once): per pixel 24 forward outputs + 16 backward inputs, 12 per covered
    pixel (texel read), vertex normals out, positions, UV indices, UVs, texture read and texture
    gradient write, position/colour gradients, SH and its gradient."""
    B, C, N, F = scene.B, scene.C, scene.N, scene.F
    P = B * C * scene.width * scene.height
    T = B * texW * texH
    return P * (24 + 16) + 12 * covered + B * C * N * 12 + B * N * 12 + 12 * F + 24 * F + 2 * 12 * T + B * N * 24 \
        + 108 * B * C


def fit_experiment(train_samples):
    """Proposal training on renders, then plain vs warm-started fit on held-out subjects."""
    import numpy as np
    import torch
    from gvv_amd import fit, proposal

    dev = torch.device("cuda", 0)
    cfg = fit.FitConfig()
    model = fit.ShapeModel(cfg, dev)
    net = proposal.ProposalNet(dev, seed=0)
    log = proposal.train(net, model, samples=train_samples, batch=16, lr=7e-4, seed=1, epochs=proposal.TRAIN_EPOCHS)
    gt = fit.sample_params(np.random.Generator(np.random.PCG64(12345)), cfg.batch, cfg.n_basis, dev)
    rep = proposal.compare_fits(net, model, gt)
    ratios = [r for r in rep.ratios if r is not None]
    return dict(metric="fit time-to-tolerance vs plain", ratio=None if rep.ratio is None else round(rep.ratio, 4),
                ratio_runs=[None if r is None else round(r, 4) for r in rep.ratios],
                ratio_spread=round(max(ratios) - min(ratios), 4) if ratios else None,
                warm_iters_runs=[it for _, it in rep.warm_runs],
                plain_runs=[dict(seconds=round(t, 4), final_loss=l) for t, l in rep.plain_runs],
                plain_seconds=round(rep.plain_seconds, 4), plain_seconds_to_final=round(rep.plain_seconds_to_final, 4),
                plain_iters=rep.plain_iters,
                plain_final_loss=rep.plain_final_loss, plain_initial_loss=rep.plain_initial_loss,
                warm_seconds=None if rep.warm_seconds is None else round(rep.warm_seconds, 4),
                warm_iters=rep.warm_iters, warm_initial_loss=rep.warm_initial_loss, fallbacks=rep.fallbacks,
                train_samples=log.samples, train_epochs=proposal.TRAIN_EPOCHS, train_seconds=round(log.seconds, 2),
                config=dict(template_vertices=model.N, template_triangles=model.F, basis=cfg.n_basis,
                            subjects=cfg.batch, views=cfg.views, resolution=cfg.resolution, adam_lr=cfg.lr))


def multi_figures(steps, warmup):
    """Batched uneven workload (32 meshes, 2k-200k triangles, 1-8 views at 256/512/1024^2)
    through the multi-mesh runtime: plain launches and one graph launch per step with default
    settings, then the autotuner and the graph again.  Returns the figures as a dict."""
    import numpy as np
    import torch
    from gvv_amd import scenes
    from gvv_amd.multi import MultiRenderer

    dev = torch.device("cuda", 0)
    batch = scenes.config2()
    mr = MultiRenderer(batch, device=dev)
    rng = np.random.Generator(np.random.PCG64(7))
    for b in mr.bufs:
        b["g_render"].copy_(torch.from_numpy(rng.normal(size=tuple(b["g_render"].shape)).astype(np.float32)))

    def timed(steps):
        for _ in range(warmup):
            mr.run(backward=True)
        torch.cuda.synchronize()
        a, z = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(steps):
            mr.run(backward=True)
        z.record()
        torch.cuda.synchronize()
        return a.elapsed_time(z) / steps

    mr.use_graph(False)
    ms_plain_default = timed(steps)
    mr.use_graph(True)
    ms_graph_default = timed(steps)
    t0 = time.perf_counter()
    mr.autotune(reps=3)
    tune_s = time.perf_counter() - t0
    ms = timed(steps)
    P = mr.pixels()
    tuning = [mr.tuning(i) for i in range(len(batch))]
    out = dict(value=round(P / (ms / 1e3) / 1e6, 2), unit="Mpix/s", ms_per_step=round(ms, 4), steps=steps,
               workload="multi32_uneven", meshes=len(batch), pixels=P, triangle_views=mr.triangles(),
               triangles_per_s=round(mr.triangles() / (ms / 1e3), 1),
               ms_plain_launches_default_tuning=round(ms_plain_default, 4),
               ms_graph_default_tuning=round(ms_graph_default, 4), autotune_seconds=round(tune_s, 2),
               tuning=[dict(mesh=s.meta["name"], tile=t, large_box=lb, size_class=c)
                       for s, (t, lb, c) in zip(batch, tuning)])
    del mr
    torch.cuda.synchronize()
    return out


def bench_multi(args):
    """--workload multi: the batched uneven workload as the line's own metric."""
    m = multi_figures(args.steps, args.warmup)
    line = {
        "metric": "fwd+bwd Mpix/s, batched uneven meshes (one graph launch per step)",
        "value": m["value"],
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (32 seeded icosphere/grid meshes)",
        "config": {k: v for k, v in m.items() if k not in ("value", "unit", "ms_per_step", "steps")},
    }
    print(json.dumps(line), flush=True)


def bench_fit(args):
    """--workload fit: the plain analysis-by-synthesis fit of the fit configuration (16 subjects x
    4 views at 512^2, 70,000-triangle template, 80 shape coefficients, masked L2, Adam), timed
    over `steps` iterations after `warmup`; one step = one iteration (shape model, render,
    loss, backward, optimizer step, the loss read on the host)."""
    import numpy as np
    import torch
    from gvv_amd import fit

    dev = torch.device("cuda", 0)
    cfg = fit.FitConfig()
    model = fit.ShapeModel(cfg, dev)
    gt = fit.sample_params(np.random.Generator(np.random.PCG64(12345)), cfg.batch, cfg.n_basis, dev)
    target, mask = fit.make_targets(model, gt)
    init = fit.plain_init(cfg.batch, cfg.n_basis, dev)
    fit.fit(model, target, mask, init, iters=max(1, args.warmup))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = fit.fit(model, target, mask, init, iters=args.steps)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    px = cfg.batch * cfg.views * cfg.resolution * cfg.resolution
    line = {
        "metric": "fit iterations/s (16 subjects x 4 views at 512^2, 70k triangles)",
        "value": round(1e3 / ms, 2),
        "unit": "iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded shape model and ground-truth parameters)",
        "config": {"workload": "fit16x4x512_vertexcolor_shaded", "subjects": cfg.batch, "views": cfg.views,
                   "resolution": cfg.resolution, "template_vertices": model.N, "template_triangles": model.F,
                   "basis": cfg.n_basis, "render_mpix_per_s": round(px / (ms / 1e3) / 1e6, 1),
                   "loss_first": res.losses[0], "loss_last": res.final_loss},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-views", type=int, default=8)
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--tuning", default=None, metavar="TILE,BOX",
                    help="fix the backward tile edge and large-face threshold instead of choosing them")
    ap.add_argument("--kernels-json", default=None, help="also write per-kernel timings here")
    ap.add_argument("--workload", choices=("render", "multi", "fit"), default="render")
    ap.add_argument("--no-fit", action="store_true", help="skip the proposal + fit time-to-tolerance run")
    ap.add_argument("--no-multi", action="store_true", help="skip the batched uneven-mesh figures")
    ap.add_argument("--train-samples", type=int, default=200_000)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MB",
                    help="after timing, run one plain device-to-device copy of MB MiB (reference kernel for "
                         "scaling rocprof FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU: every (batch, view) pair renders independently and the renderer "
              f"runs on a single MI355X; --gpus {args.gpus} (WORLD_SIZE={world}) is not supported.",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "multi":
        return bench_multi(args)
    if args.workload == "fit":
        return bench_fit(args)

    import numpy as np
    import torch
    from gvv_amd import Rasterizer, scenes

    dev = torch.device("cuda", 0)
    scene = scenes.config1()
    texH, texW = scene.tex.shape[1], scene.tex.shape[2]
    rast = Rasterizer(**scene.renderer_kwargs())
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    vpos, vcol, tex, sh, extr, intr = (t(a) for a in scene.inputs())
    target = t(scene.target)

    bary, face, render, vnormal, _ = rast.forward(vpos, vcol, tex, sh, extr, intr)
    g_render = (2.0 * (render - target)).contiguous()
    covered = int((face >= 0).sum().item())

    def step():
        _, f, _, vn, _ = rast.forward(vpos, vcol, tex, sh, extr, intr)
        rast.backward(g_render, None, None, vpos, vcol, tex, sh, vn, f, extr, intr)

    # launch settings: the fastest of the autotuner's candidates on this workload (untimed), or
    # the ones given with --tuning
    tuning_ms = {}
    candidates = [(t, b) for t in (8, 16, 32) for b in (64, 256, 1024)]
    if args.tuning:
        candidates = [tuple(int(v) for v in args.tuning.split(","))]
    for tile, box in candidates:
        rast.set_tuning(tile, box)
        step()
        a, z = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record()
        for _ in range(3):
            step()
        z.record()
        torch.cuda.synchronize()
        tuning_ms[(tile, box)] = a.elapsed_time(z) / 3
    best = min(tuning_ms, key=tuning_ms.get)
    # the library's defaults (tile 16, large box 256) stay unless a candidate beats them by more
    # than the 3-step trials' noise (about 1%): a toss-up then runs at the settings the committed
    # counter profiles were recorded at
    if (16, 256) in tuning_ms and tuning_ms[best] > 0.99 * tuning_ms[(16, 256)]:
        best = (16, 256)
    rast.set_tuning(*best)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    def timed(steps):
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        start.record()
        for _ in range(steps):
            step()
        end.record()
        torch.cuda.synchronize()
        return start.elapsed_time(end) / steps

    # the throughput: K steps with nothing else on the stream
    ms = timed(args.steps)
    # the same K steps again with the library's per-kernel HIP events on its launch stream, for
    # the per-kernel figures (the events themselves add a little to this second region)
    rast.profile(True)
    ms_instrumented = timed(args.steps)
    kernels = rast.profile_report()
    rast.profile(False)

    P = scene.B * scene.C * scene.width * scene.height
    value = P / (ms / 1e3) / 1e6
    dom = max(kernels, key=lambda k: kernels[k][0])
    dom_ms = kernels[dom][0] / kernels[dom][1]
    dom_bytes = kernel_bytes(dom, scene, covered, texW, texH)
    workload = "cylinder100k_8x1024"
    counters = load_counters(workload)
    stamp = None
    stale_for = lambda k: False
    if counters is not None:
        rec = counters.get("recorded", {})
        digest = source_digest()
        rec_tile, rec_box = (list(rec.get("tuning", [])) + [None, None])[:2]
        # the backward kernels depend on the tile edge only, the depth pass on the large-face
        # threshold only; other sources make every figure stale
        sources_stale = rec.get("sources") != digest
        tile_stale, box_stale = rec_tile != best[0], rec_box != best[1]
        stale_for = lambda k: sources_stale or (tile_stale if k in ("k_pixel_grad", "k_grad_finish") else box_stale)
        if sources_stale:
            print(f"bench.py: counter figures in profiles/counters_{workload}.json are stale: recorded from kernel "
                  f"sources {rec.get('sources')}, these are {digest}", file=sys.stderr, flush=True)
        elif tile_stale or box_stale:
            print(f"bench.py: counter figures in profiles/counters_{workload}.json were recorded at tuning "
                  f"{rec.get('tuning')}, this run chose {list(best)}", file=sys.stderr, flush=True)
        stamp = dict(file=f"profiles/counters_{workload}.json", tuning=rec.get("tuning"), commit=rec.get("commit"),
                     sources=rec.get("sources"), stale=stale_for(dom))
    roofline = None
    if dom_bytes is not None:
        achieved = dom_bytes / (dom_ms / 1e3) / 1e9
        traffic = insts = None
        if counters is not None:
            traffic = counters.get("traffic_bytes", {}).get(dom)
            insts = counters.get("valu_insts", {}).get(dom)
        hbm = dict(achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4),
                   traffic=traffic, algorithmic_bytes=dom_bytes,
                   traffic_frac=None if traffic is None else round(traffic / (dom_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4))
        valu = None
        if insts:
            rate = insts / (dom_ms / 1e3)
            valu = dict(achieved=round(rate / 1e9, 1), peak=round(VALU_PEAK_PER_S / 1e9, 1), unit="Ginst/s",
                        frac=round(rate / VALU_PEAK_PER_S, 4), insts=insts)
        # the bound is the resource the kernel uses the larger share of: HBM (counter traffic, or
        # the algorithmic bytes when no traffic was recorded) or VALU instruction issue
        hbm_use = hbm["traffic_frac"] if hbm["traffic_frac"] is not None else hbm["frac"]
        top = valu if valu is not None and valu["frac"] > hbm_use else hbm
        roofline = dict(bound="valu_issue" if top is valu else "hbm", kernel=dom, achieved=top["achieved"],
                        peak=top["peak"], unit=top["unit"], frac=top["frac"], traffic=traffic, avg_ms=round(dom_ms, 4),
                        hbm=hbm, valu=valu, counters=stamp)

    atomics = None
    if counters is not None and counters.get("atomic_requests"):
        counts = counters["atomic_requests"]
        peak = atomic_peak(dev)
        atomics = {}
        for k in ("k_pixel_grad", "k_raster", "k_raster_large", "k_grad_finish"):
            if k in counts and k in kernels:
                rate = counts[k] / (kernels[k][0] / kernels[k][1] / 1e3)
                atomics[k] = dict(requests=counts[k], rate_per_s=round(rate, 1), peak_per_s=round(peak["streaming"], 1),
                                  frac=round(rate / peak["streaming"], 4), stale=stale_for(k))
        atomics["peak_scattered_per_s"] = round(peak["scattered"], 1)
        # the reference formulation scatters ~210 atomics per covered pixel in textured+shaded mode
        # (3 texel + 27 SH + 9 barycentric-position + 9 model-to-data + ~162 per-incident-face normal terms)
        atomics["reference_formulation_per_step"] = 210 * covered
        atomics["issued_requests_per_step"] = sum(v for k, v in counts.items() if k in kernels)
        atomics["counters"] = stamp

    if args.calibrate_copy > 0:
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # the outputs of one step at the chosen settings, for the accuracy check against the oracle
    gb, gf, gr, gvn, _ = rast.forward(vpos, vcol, tex, sh, extr, intr)
    gg = rast.backward(g_render, None, None, vpos, vcol, tex, sh, gvn, gf, extr, intr)
    torch.cuda.synchronize()
    gpu_out = dict(bary=gb.cpu().numpy(), face=gf.cpu().numpy(), render=gr.cpu().numpy(), vnormal=gvn.cpu().numpy(),
                   **{"g_" + k: v.cpu().numpy() for k, v in zip(("pos", "col", "tex", "sh"), gg)})
    g_render_np = g_render.cpu().numpy()
    del gb, gf, gr, gvn, gg

    cpu = accuracy = None
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu, (ref, gref, full) = cpu_baseline(scene, threads, views=args.cpu_views, reps=args.cpu_reps,
                                              g_render=g_render_np)
        if full:
            accuracy = accuracy_report(gpu_out, ref, gref)
        del ref, gref
    del gpu_out

    multi = None
    if not args.no_multi:
        try:
            multi = multi_figures(args.steps, args.warmup)
        except Exception as e:      # reported in the line; the throughput figure stands on its own
            multi = dict(error=f"{type(e).__name__}: {e}")

    fit = None
    if not args.no_fit:
        del rast
        torch.cuda.synchronize()
        try:
            fit = fit_experiment(args.train_samples)
        except Exception as e:      # reported in the line; the throughput figure stands on its own
            fit = dict(error=f"{type(e).__name__}: {e}")

    sb = step_bytes(scene, covered, texW, texH)
    step_roofline = dict(algorithmic_bytes=sb, achieved_GBs=round(sb / (ms / 1e3) / 1e9, 1), peak_GBs=HBM_PEAK_GBS,
                         frac=round(sb / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4))

    line = {
        "metric": "fwd+bwd Mpix/s (8x1024^2 views, 100k tris)",
        "value": round(value, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded capped cylinder, smooth-noise texture, ring cameras)",
        "ms_per_step_with_kernel_events": round(ms_instrumented, 4),
        "config": {"workload": "cylinder100k_8x1024_textured_shaded", "triangles": scene.F,
                   "vertices": scene.N, "views": scene.C, "resolution": [scene.height, scene.width],
                   "texture": [texH, texW], "modes": "textured+shaded", "covered_px": covered,
                   "triangles_per_s": round(scene.B * scene.C * scene.F / (ms / 1e3), 1),
                   "tuning": {"tile": best[0], "large_box": best[1],
                              "candidates_ms": {f"{t}/{b}": round(v, 4) for (t, b), v in tuning_ms.items()}}},
        "roofline": roofline,
        "step_roofline": step_roofline,
        "atomics": atomics,
        "cpu_baseline": cpu,
        "kernels_ms": {k: round(v[0] / v[1], 4) for k, v in kernels.items()},
        "accuracy": accuracy,
        "fit": fit,
        "multi": multi,
    }
    if args.kernels_json:
        with open(args.kernels_json, "w") as fh:
            json.dump(dict(kernels=kernels, ms_per_step=ms), fh, indent=1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
