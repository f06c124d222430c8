#!/usr/bin/env python3
"""Throughput of the differentiable rasterizer, forward + backward, on one MI355X.

Workload: the capped cylinder of 100,000 triangles / 50,002 vertices, 1024^2 smooth-noise
texture, 8 views at 1024x1024 on a 3.5 m ring, textured + shaded, full backward to vertices,
texture and SH.  One step = one forward and one backward of the renderer over the 8 views with
every input already resident in HBM; the upstream gradient is the sum-of-squares loss gradient
2(render - target) of the first forward, fixed across steps.

Prints one JSON line: Mpix/s = 8 x 1024 x 1024 / (t_fwd + t_bwd), plus the roofline of the
dominant kernel (HIP-event timing through the library's profiling hook) and the CPU oracle
timed on a bounded sample of the same workload.
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


def cpu_baseline(scene, threads, views=1):
    """Oracle forward + backward on `views` of the 8 views (bounded sample), multithreaded."""
    import numpy as np
    from oracle import oracle as O
    O.set_threads(threads)
    kw = scene.renderer_kwargs()
    kw["n_cameras"] = views
    orc = O.Oracle(**kw)
    ins = (scene.vpos, scene.vcol, scene.tex, np.ascontiguousarray(scene.sh[:, :views]),
           np.ascontiguousarray(scene.extr[:, :views]), np.ascontiguousarray(scene.intr[:, :views]))
    t0 = time.perf_counter()
    out = orc.forward(*ins)
    t1 = time.perf_counter()
    target = np.ascontiguousarray(scene.target[:, :views])
    g = (2.0 * (out["render"] - target)).astype(np.float32)
    t2 = time.perf_counter()
    orc.backward(g, np.zeros_like(g), target, scene.vpos, scene.vcol, scene.tex, ins[3], out["vnormal"], out["bary"],
                 out["face"], ins[4], ins[5])
    t3 = time.perf_counter()
    px = views * scene.width * scene.height * scene.B
    secs = (t1 - t0) + (t3 - t2)
    return dict(value=px / secs / 1e6, unit="Mpix/s", cores=threads, kind="port",
                sample=f"FP64 oracle fwd+bwd on {views} of {scene.C} views ({px} px, {scene.F} triangles); "
                       f"fwd {t1 - t0:.2f}s bwd {t3 - t2:.2f}s")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-views", type=int, default=2)
    ap.add_argument("--kernels-json", default=None, help="also write per-kernel timings here")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MB",
                    help="after timing, run one plain device-to-device copy of MB MiB (reference kernel for "
                         "scaling rocprof FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU: every (batch, view) pair renders independently and this build "
              f"runs on a single MI355X; --gpus {args.gpus} (WORLD_SIZE={world}) is not supported.",
              file=sys.stderr)
        sys.exit(2)

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

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    rast.profile(True)
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(args.steps):
        step()
    end.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / args.steps
    kernels = rast.profile_report()
    rast.profile(False)

    P = scene.B * scene.C * scene.width * scene.height
    value = P / (ms / 1e3) / 1e6
    dom = max(kernels, key=lambda k: kernels[k][0])
    dom_ms = kernels[dom][0] / kernels[dom][1]
    dom_bytes = kernel_bytes(dom, scene, covered, texW, texH)
    roofline = None
    if dom_bytes is not None:
        achieved = dom_bytes / (dom_ms / 1e3) / 1e9
        traffic = None
        tf = os.path.join(ROOT, "profiles", "traffic.json")
        if os.path.exists(tf):
            with open(tf) as fh:
                traffic = json.load(fh).get(dom)
        roofline = dict(bound="hbm", kernel=dom, achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                        frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, algorithmic_bytes=dom_bytes,
                        avg_ms=round(dom_ms, 4))

    if args.calibrate_copy > 0:
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    cpu = None
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = cpu_baseline(scene, threads, views=args.cpu_views)

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
        "config": {"workload": "cylinder100k_8x1024_textured_shaded", "triangles": scene.F,
                   "vertices": scene.N, "views": scene.C, "resolution": [scene.height, scene.width],
                   "texture": [texH, texW], "modes": "textured+shaded", "covered_px": covered,
                   "triangles_per_s": round(scene.B * scene.C * scene.F / (ms / 1e3), 1)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels_ms": {k: round(v[0] / v[1], 4) for k, v in kernels.items()},
    }
    if args.kernels_json:
        with open(args.kernels_json, "w") as fh:
            json.dump(dict(kernels=kernels, ms_per_step=ms), fh, indent=1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
