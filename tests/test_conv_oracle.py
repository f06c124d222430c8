"""Pins of the FP64 conv oracle (oracle/conv_oracle.py) against torch's float64 CPU conv2d and
its autograd, and of the bf16 rounding helper against torch's bfloat16 conversion."""
import numpy as np
import pytest
import torch

from oracle import conv_oracle as CO


@pytest.mark.parametrize("B,H,Cin,Cout,ks,st,pad", [(2, 9, 3, 5, 3, 2, 1), (1, 8, 4, 6, 3, 1, 1), (2, 11, 2, 3, 5, 2, 2),
                                                     (3, 1, 16, 7, 1, 1, 0), (1, 7, 2, 2, 3, 3, 0)])
def test_conv_oracle_matches_torch(B, H, Cin, Cout, ks, st, pad):
    rng = np.random.default_rng(B * 100 + H)
    x = rng.standard_normal((B, H, H, Cin))
    w = rng.standard_normal((Cout, ks, ks, Cin))
    b = rng.standard_normal(Cout)
    tx = torch.tensor(x.transpose(0, 3, 1, 2), dtype=torch.float64, requires_grad=True)
    tw = torch.tensor(w.transpose(0, 3, 1, 2), dtype=torch.float64, requires_grad=True)
    tb = torch.tensor(b, dtype=torch.float64, requires_grad=True)
    ty = torch.nn.functional.conv2d(tx, tw, tb, stride=st, padding=pad)
    y = CO.conv_forward(x, w, b, st, pad)
    assert y.shape == (B, CO.out_size(H, ks, st, pad), CO.out_size(H, ks, st, pad), Cout)
    np.testing.assert_allclose(y, ty.detach().numpy().transpose(0, 2, 3, 1), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(CO.conv_forward(x, w, b, st, pad, relu=True), np.maximum(y, 0), rtol=0, atol=0)
    dz = rng.standard_normal(y.shape)
    ty.backward(torch.tensor(dz.transpose(0, 3, 1, 2)))
    dx, dw, db = CO.conv_backward(x, w, dz, st, pad)
    np.testing.assert_allclose(dx, tx.grad.numpy().transpose(0, 2, 3, 1), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(dw, tw.grad.numpy().transpose(0, 2, 3, 1), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(db, tb.grad.numpy(), rtol=1e-12, atol=1e-12)


def test_abs_bounds_dominate():
    rng = np.random.default_rng(5)
    x, w, dz = rng.standard_normal((2, 6, 6, 3)), rng.standard_normal((4, 3, 3, 3)), rng.standard_normal((2, 3, 3, 4))
    y = CO.conv_forward(x, w, np.zeros(4), 2, 1)
    assert np.all(CO.conv_abs_bound(x, w, 2, 1) >= np.abs(y))
    for a, bnd in zip(CO.conv_backward(x, w, dz, 2, 1), CO.conv_backward_abs_bound(x, w, dz, 2, 1)):
        assert np.all(bnd >= np.abs(a) - 1e-12)


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(1)
    a = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 10000),
                        np.array([0.0, -0.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 3.4e38, np.inf, -np.inf], np.float32)])
    want = torch.from_numpy(a).to(torch.bfloat16).float().numpy()
    np.testing.assert_array_equal(CO.bf16_round(a), want)
    np.testing.assert_array_equal(CO.bits_to_f32(CO.bf16_bits(want)), want)
