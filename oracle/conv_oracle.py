"""FP64 CPU oracle of the proposal network's convolutions (test infrastructure only).

The reference has no learned component, so there is no reference file to follow: this is
the textbook strided 2-D cross-correlation in NHWC layout, with weights [Cout][k][k][Cin],
zero padding, and its two adjoints.  It is pinned in tests against torch's float64 CPU
conv2d and autograd.  Only tests, the smoke check and the bench's CPU baseline use it.
"""
from __future__ import annotations

import numpy as np


def bf16_round(a):
    """float32 -> nearest bfloat16 (ties to even), returned as float32 values."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    out = u.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), a, out)


def bf16_bits(a):
    """float32 values exactly representable in bfloat16 -> their uint16 bit patterns."""
    return (np.ascontiguousarray(a, dtype=np.float32).view(np.uint32) >> 16).astype(np.uint16)


def bits_to_f32(h):
    return (np.asarray(h, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def out_size(n, k, stride, pad):
    return (n + 2 * pad - k) // stride + 1


def _patches(x, k, stride, pad):
    """x [B,H,W,C] -> [B,OH,OW,k,k,C] (zero padded)."""
    B, H, W, C = x.shape
    xp = np.zeros((B, H + 2 * pad, W + 2 * pad, C), dtype=np.float64)
    xp[:, pad:pad + H, pad:pad + W] = x
    OH, OW = out_size(H, k, stride, pad), out_size(W, k, stride, pad)
    win = np.lib.stride_tricks.sliding_window_view(xp, (k, k), axis=(1, 2))    # B, H', W', C, k, k
    win = win[:, 0:(OH - 1) * stride + 1:stride, 0:(OW - 1) * stride + 1:stride]
    return np.transpose(win, (0, 1, 2, 4, 5, 3))


def conv_forward(x, w, b, stride, pad, relu=False):
    """y[b,oh,ow,o] = sum_{i,j,c} x[b, oh*s-p+i, ow*s-p+j, c] w[o,i,j,c] + b[o]; optional ReLU."""
    cols = _patches(np.asarray(x, np.float64), w.shape[1], stride, pad)
    y = np.einsum("bhwijc,oijc->bhwo", cols, np.asarray(w, np.float64), optimize=True) + np.asarray(b, np.float64)
    return np.maximum(y, 0.0) if relu else y


def conv_abs_bound(x, w, stride, pad):
    """sum_{i,j,c} |x||w| per output element: the scale of the accumulation error."""
    return conv_forward(np.abs(x), np.abs(w), np.zeros(w.shape[0]), stride, pad)


def conv_backward(x, w, dz, stride, pad):
    """Adjoints of conv_forward (without the activation) at output gradient dz:
    returns dx [B,H,W,Cin], dw [Cout,k,k,Cin], db [Cout]."""
    x = np.asarray(x, np.float64)
    w = np.asarray(w, np.float64)
    dz = np.asarray(dz, np.float64)
    k = w.shape[1]
    cols = _patches(x, k, stride, pad)
    dw = np.einsum("bhwo,bhwijc->oijc", dz, cols, optimize=True)
    db = dz.sum((0, 1, 2))
    dcols = np.einsum("bhwo,oijc->bhwijc", dz, w, optimize=True)
    B, H, W, C = x.shape
    OH, OW = dz.shape[1], dz.shape[2]
    dxp = np.zeros((B, H + 2 * pad, W + 2 * pad, C))
    for i in range(k):
        for j in range(k):
            dxp[:, i:i + (OH - 1) * stride + 1:stride, j:j + (OW - 1) * stride + 1:stride] += dcols[:, :, :, i, j]
    return dxp[:, pad:pad + H, pad:pad + W], dw, db


def conv_backward_abs_bound(x, w, dz, stride, pad):
    """Scales |.|-sums of the three adjoints (accumulation-error scales)."""
    return conv_backward(np.abs(x), np.abs(w), np.abs(dz), stride, pad)
