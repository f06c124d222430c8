"""BF16 MFMA convolution kernels (include/gvv_proposal.h) against the FP64 conv oracle.

Inputs are drawn in float32 and rounded to bf16; the oracle runs in FP64 on the same rounded
values.  Bounds per element: bf16 outputs within half a bf16 ulp of the exact value plus the
FP32 accumulation error, fp32 outputs within the accumulation error, where the accumulation
error is 2e-5 x the sum of |terms| (conv_abs_bound).
"""
import ctypes

import numpy as np
import pytest
import torch

from gvv_amd import _lib
from oracle import conv_oracle as CO

pytestmark = pytest.mark.gpu

ACC = 2e-5          # fp32 accumulation error per unit of sum |terms|
HALF_ULP = 2.0 ** -8  # bf16 rounding of the result (relative, round to nearest)


def P(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def S():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def bf(a, device):
    """float32 numpy (bf16-representable) -> bf16 device tensor with the same values."""
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(device).to(torch.bfloat16)


def rnd(rng, shape, scale=1.0):
    return CO.bf16_round(rng.standard_normal(shape).astype(np.float32) * scale)


def check_close(name, got, want, bound, bf16_out):
    tol = bound * ACC + (np.abs(want) * HALF_ULP if bf16_out else 0.0) + 1e-30
    err = np.abs(got.astype(np.float64) - want)
    worst = float((err / tol).max()) if err.size else 0.0
    rel = float(np.linalg.norm(err) / max(np.linalg.norm(want), 1e-30))
    print(f"{name}: worst err/bound {worst:.3f}  normwise rel {rel:.2e}", flush=True)
    assert worst <= 1.0, name


GEOMS = [
    # B, H, Cin, Cout, ks, stride, pad
    (2, 32, 8, 32, 3, 2, 1),        # first layer shape (K = 72: partial k step)
    (2, 16, 32, 64, 3, 2, 1),
    (3, 9, 16, 24, 3, 2, 1),        # odd sizes: ragged pixel and channel tiles
    (2, 8, 64, 128, 3, 2, 1),
    (1, 6, 128, 256, 3, 2, 1),
    (4, 1, 512, 128, 1, 1, 0),      # linear head as a 1x1 convolution
    (2, 7, 8, 16, 3, 1, 1),         # stride 1
    (2, 10, 16, 8, 5, 2, 2),        # 5x5 kernel
]


@pytest.mark.parametrize("geom", GEOMS, ids=[f"B{g[0]}_H{g[1]}_{g[2]}to{g[3]}_k{g[4]}s{g[5]}" for g in GEOMS])
def test_conv_forward_backward_vs_oracle(device, geom):
    B, H, Cin, Cout, ks, st, pad = geom
    L = _lib.load()
    rng = np.random.default_rng(hash(geom) % 2**32)
    OH = CO.out_size(H, ks, st, pad)
    x = np.maximum(rnd(rng, (B, H, H, Cin)), 0.0)                 # a ReLU output (as in the network)
    w = rnd(rng, (Cout, ks, ks, Cin), 1.0 / np.sqrt(ks * ks * Cin))
    b = rng.standard_normal(Cout).astype(np.float32) * 0.1
    dz = rnd(rng, (B, OH, OH, Cout))

    xd, dzd = bf(x, device), bf(dz, device)
    wf = torch.from_numpy(w).to(device)
    wb = torch.empty(w.shape, dtype=torch.bfloat16, device=device)
    wt = torch.empty((Cin, ks, ks, Cout), dtype=torch.bfloat16, device=device)
    _lib.check(L.gvv_conv_prepare_weights(Cout, ks * ks, Cin, P(wf), P(wb), P(wt), S()), "prepare")
    assert torch.equal(wb.float().cpu(), torch.from_numpy(w))
    assert torch.equal(wt.float().cpu(), torch.from_numpy(np.ascontiguousarray(np.transpose(w, (3, 1, 2, 0)))))
    bd = torch.from_numpy(b).to(device)

    # forward, bf16 output with ReLU, and fp32 output without
    y = torch.empty((B, OH, OH, Cout), dtype=torch.bfloat16, device=device)
    _lib.check(L.gvv_conv_forward(B, H, H, Cin, Cout, ks, st, pad, P(xd), P(wb), P(bd), 1, P(y), 0, S()), "fwd")
    yf = torch.empty((B, OH, OH, Cout), dtype=torch.float32, device=device)
    _lib.check(L.gvv_conv_forward(B, H, H, Cin, Cout, ks, st, pad, P(xd), P(wb), P(bd), 0, P(yf), 1, S()), "fwd f32")
    ref = CO.conv_forward(x, w, b, st, pad)
    bound = CO.conv_abs_bound(x, w, st, pad) + np.abs(b)
    check_close("forward f32", yf.cpu().numpy(), ref, bound, False)
    check_close("forward relu bf16", y.float().cpu().numpy(), np.maximum(ref, 0.0), bound, True)

    # backward data, with and without the ReLU mask of this layer's input
    mask_src = rnd(rng, (B, H, H, Cin))                            # signs decide the mask
    md = bf(mask_src, device)
    dx = torch.empty((B, H, H, Cin), dtype=torch.bfloat16, device=device)
    dxm = torch.empty_like(dx)
    _lib.check(L.gvv_conv_backward_data(B, H, H, Cin, Cout, ks, st, pad, P(dzd), P(wt), None, P(dx), S()), "bwd data")
    _lib.check(L.gvv_conv_backward_data(B, H, H, Cin, Cout, ks, st, pad, P(dzd), P(wt), P(md), P(dxm), S()), "bwd data mask")
    rdx, rdw, rdb = CO.conv_backward(x, w, dz, st, pad)
    bdx, bdw, bdb = CO.conv_backward_abs_bound(x, w, dz, st, pad)
    check_close("backward data", dx.float().cpu().numpy(), rdx, bdx, True)
    check_close("backward data masked", dxm.float().cpu().numpy(), rdx * (mask_src > 0), bdx, True)

    # backward weight + bias, twice (the result is overwritten, not accumulated; deterministic)
    need = L.gvv_conv_workspace_floats(B, H, H, Cin, Cout, ks, st, pad)
    ws = torch.empty(int(need), device=device)
    dw = torch.full((Cout, ks, ks, Cin), 7.0, device=device)
    db = torch.full((Cout,), 7.0, device=device)
    for it in range(2):
        _lib.check(L.gvv_conv_backward_weight(B, H, H, Cin, Cout, ks, st, pad, P(xd), P(dzd), P(dw), P(db), P(ws),
                                              ws.numel(), S()), "bwd weight")
        if it == 0:
            first = (dw.clone(), db.clone())
    assert torch.equal(first[0], dw) and torch.equal(first[1], db)
    check_close("backward weight", dw.cpu().numpy(), rdw, bdw, False)
    check_close("backward bias", db.cpu().numpy(), rdb, bdb, False)


@pytest.mark.parametrize("features", [0, 1])
def test_pack_mosaic(device, features):
    L = _lib.load()
    rng = np.random.default_rng(0)
    B, V, R, d = 2, 3, 8, 2
    render = rng.random((B, V, R, R, 3)).astype(np.float32)
    render[0, 1, :3, :5] = (0.0, 1.0, 0.0)          # background pixels
    rd = torch.from_numpy(render).to(device)
    cols = 2
    x = torch.full((B, 2 * R // d, 2 * R // d, 8), 5.0, dtype=torch.bfloat16, device=device)
    _lib.check(L.gvv_pack_mosaic(P(rd), B, V, R, R, cols, d, features, P(x), S()), "pack")
    got = x.float().cpu().numpy()
    want = np.zeros((B, 2 * R // d, 2 * R // d, 8), np.float32)
    r = R // d
    coord = (2.0 * (np.arange(r) + 0.5) / r - 1.0).astype(np.float32)
    for v in range(V):
        ty, tx = divmod(v, cols)
        blocks = render[:, v].reshape(B, r, d, r, d, 3)
        pooled = blocks.mean((2, 4)).astype(np.float32)
        tile = want[:, ty * r:(ty + 1) * r, tx * r:(tx + 1) * r]
        if features:
            tile[..., :3] = CO.bf16_round(pooled - 0.5)
            bg = np.all(blocks == np.array([0.0, 1.0, 0.0], np.float32), axis=-1)
            tile[..., 3] = CO.bf16_round((~bg).mean((2, 4)).astype(np.float32))
            tile[..., 4] = CO.bf16_round(np.broadcast_to(coord[None, None, :], (B, r, r)))
            tile[..., 5] = CO.bf16_round(np.broadcast_to(coord[None, :, None], (B, r, r)))
        else:
            tile[..., :3] = CO.bf16_round(pooled)
    # features: one bf16 ulp of slack for the order of the block sum (values in [-0.5, 1])
    np.testing.assert_allclose(got, want, rtol=0, atol=2.0 ** -8 if features else 0)
