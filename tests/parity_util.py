"""Shared GPU-vs-oracle comparison used by the GPU test modules.

Accuracy policy (per output tensor, against the oracle on the same inputs):
  * face ids, barycentrics and vertex normals: bit-exact;
  * render and normal map: max-norm relative error max|x - ref| / max|ref| <= 1e-5, and
    elementwise |x - ref| <= 1e-5 |ref| + 1e-6 max|ref| for every entry.  The absolute part
    covers values that are themselves a cancellation (the normal mode's (1 + n) / 2 near
    n = -1, an SH sum near zero), where a binary32 evaluation of the same formula is off by
    its rounding of the terms, not of the result;
  * gradients: max-norm relative error <= 1e-4, and over the entries with
    |ref| >= 1e-3 max|ref| the elementwise relative error has median <= 1e-5 and 99.9th
    percentile <= 1e-3.  Gradient entries are sums over many pixels whose terms can cancel;
    where they do, a float32 sum (and the order-dependent float atomics that form it) has a
    relative error set by the cancelled magnitude, not by the entry's own, so a few entries of
    a tensor can exceed any fixed elementwise bound without any term being wrong (the largest
    seen is 3e-4).  The median and percentile bounds catch systematic errors (a dropped or
    mis-signed term moves most entries); the max-norm bound catches any large absolute error.
    Across the suite the measured figures are at most 4e-6 (max-norm), 6e-5 (99.9th
    percentile) and 6e-7 (median).
Every comparison is recorded in ACCURACY; the session summary (conftest) prints the table.
"""
import numpy as np
import torch

from conftest import rel_err
from gvv_amd import Rasterizer
from oracle import oracle as O

RENDER_TOL = 1e-5
GRAD_TOL = 1e-4
ELEM_RENDER_RTOL = 1e-5
ELEM_RENDER_ATOL = 1e-6         # times max|ref|
ELEM_GRAD_MEDIAN = 1e-5
ELEM_GRAD_P999 = 1e-3
ELEM_FLOOR = 1e-3

ACCURACY = []      # dicts: case, output, maxnorm, elem_max, elem_p999, elem_median, n


def to_dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def accuracy(case, name, x, ref, floor=ELEM_FLOOR):
    """Max-norm and elementwise relative errors of x against ref; recorded in ACCURACY."""
    x = np.asarray(x, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    rec = dict(case=case, output=name, maxnorm=rel_err(x, ref), elem_max=0.0, elem_p999=0.0, elem_median=0.0, n=0)
    amax = np.abs(ref).max() if ref.size else 0.0
    if amax > 0:
        m = np.abs(ref) >= floor * amax
        r = np.abs(x[m] - ref[m]) / np.abs(ref[m])
        rec.update(elem_max=float(r.max()), elem_p999=float(np.quantile(r, 0.999)), elem_median=float(np.median(r)),
                   n=int(m.sum()))
    ACCURACY.append(rec)
    print(f"    {name:7s} maxnorm {rec['maxnorm']:.2e}  elementwise max {rec['elem_max']:.2e} p99.9 "
          f"{rec['elem_p999']:.2e} median {rec['elem_median']:.2e} over {rec['n']} entries", flush=True)
    return rec


def check_render(case, name, x, ref):
    rec = accuracy(case, name, x, ref)
    assert rec["maxnorm"] <= RENDER_TOL, f"{case} {name}: max-norm {rec['maxnorm']:.3e}"
    x = np.asarray(x, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    if ref.size:
        bound = ELEM_RENDER_RTOL * np.abs(ref) + ELEM_RENDER_ATOL * np.abs(ref).max()
        bad = int((np.abs(x - ref) > bound).sum())
        assert bad == 0, f"{case} {name}: {bad} entries beyond 1e-5 |ref| + 1e-6 max|ref|"


def check_grad(case, name, x, ref):
    rec = accuracy(case, name, x, ref)
    assert rec["maxnorm"] <= GRAD_TOL, f"{case} grad {name}: max-norm {rec['maxnorm']:.3e}"
    assert rec["elem_median"] <= ELEM_GRAD_MEDIAN, f"{case} grad {name}: elementwise median {rec['elem_median']:.3e}"
    assert rec["elem_p999"] <= ELEM_GRAD_P999, f"{case} grad {name}: elementwise p99.9 {rec['elem_p999']:.3e}"


def check_exact(case, face, bary, vnormal, ref):
    n_bad = int((face != ref["face"]).sum())
    ACCURACY.append(dict(case=case, output="face", maxnorm=float(n_bad), elem_max=0.0, elem_p999=0.0, elem_median=0.0,
                         n=int(face.size)))
    assert n_bad == 0, f"{case}: face ids differ at {n_bad} px"
    assert np.array_equal(bary.view(np.uint32), ref["bary"].view(np.uint32)), f"{case}: barycentrics not bit-exact"
    assert np.array_equal(vnormal.view(np.uint32), ref["vnormal"].view(np.uint32)), f"{case}: vertex normals not bit-exact"


def run_parity(scene, device, albedo, shading, m2d, seed=0, image_fs=1, use_bary=False, check_grads=True, tuning=None,
               case=None, depth_pass=None):
    kw = scene.renderer_kwargs()
    kw.update(albedo_mode=albedo, shading_mode=shading)
    case = case or f"{scene.meta.get('name')} {albedo}/{shading}{' m2d' if m2d else ''}"
    orc = O.Oracle(**kw, image_filter_size=image_fs)
    ref = orc.forward(*scene.inputs())
    rast = Rasterizer(**kw, image_filter_size=image_fs)
    if tuning is not None:
        rast.set_tuning(*tuning)
        assert rast.tuning() == tuple(tuning)
    if depth_pass is not None:
        rast.set_depth_pass(depth_pass)
        assert rast.depth_pass() == depth_pass
    ins = [to_dev(a, device) for a in scene.inputs()]
    bary, face, render, vnormal, _ = rast.forward(*ins)
    torch.cuda.synchronize()
    face_np = face.cpu().numpy()
    print(f"[{case}] covered px {int((face_np >= 0).sum())}", flush=True)
    check_exact(case, face_np, bary.cpu().numpy(), vnormal.cpu().numpy(), ref)
    check_render(case, "render", render.cpu().numpy(), ref["render"])
    if not check_grads:
        return
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    g_render = rng.normal(size=ref["render"].shape).astype(np.float32)
    g_target = rng.normal(size=ref["render"].shape).astype(np.float32) if m2d else None
    gref = orc.backward(g_render, g_target if m2d else np.zeros_like(g_render), scene.target, scene.vpos, scene.vcol,
                        scene.tex, scene.sh, ref["vnormal"], ref["bary"], ref["face"], scene.extr, scene.intr)
    got = rast.backward(to_dev(g_render, device), to_dev(g_target, device) if m2d else None, to_dev(scene.target, device),
                        ins[0], ins[1], ins[2], ins[3], vnormal, face, ins[4], ins[5], bary=bary if use_bary else None)
    torch.cuda.synchronize()
    for name, g in zip(("pos", "col", "tex", "sh"), got):
        check_grad(case, "g_" + name, g.cpu().numpy(), gref[name])
