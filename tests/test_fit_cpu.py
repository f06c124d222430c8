"""CPU checks of the fitting harness's parameter model (no renderer involved)."""
import math

import numpy as np
import torch

from gvv_amd import fit, scenes


def test_rodrigues_closed_forms():
    w = torch.tensor([[0.0, 0.0, 0.0], [0.0, 0.0, math.pi / 2], [0.3, -0.2, 0.5], [1e-6, 0.0, 0.0]], dtype=torch.float64)
    R = fit.rodrigues(w)
    assert torch.allclose(R[0], torch.eye(3, dtype=torch.float64))
    assert torch.allclose(R[1], torch.tensor([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]], dtype=torch.float64), atol=1e-12)
    for r in R:
        assert torch.allclose(r @ r.T, torch.eye(3, dtype=torch.float64), atol=1e-12)
        assert abs(torch.det(r).item() - 1.0) < 1e-12
    # rotation about its own axis leaves the axis fixed
    assert torch.allclose(R[2] @ w[2], w[2], atol=1e-12)
    # derivative at zero is the cross-product matrix (finite difference)
    h = 1e-6
    for k in range(3):
        e = torch.zeros(1, 3, dtype=torch.float64)
        e[0, k] = h
        dR = (fit.rodrigues(e) - fit.rodrigues(-e))[0] / (2 * h)
        K = torch.zeros(3, 3, dtype=torch.float64)
        ax = torch.zeros(3, dtype=torch.float64)
        ax[k] = 1.0
        K = torch.tensor([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]], dtype=torch.float64)
        assert torch.allclose(dR, K, atol=1e-8)


def test_shape_basis_orthonormal_and_scaled():
    rng = np.random.Generator(np.random.PCG64(0))
    v, f, _ = scenes.capped_cylinder(40, 30, 1.0, 2.0, rng, 0.05)
    X = fit._smooth_fields(rng, f, v.shape[0], 8).reshape(8, -1).T
    Q, _ = np.linalg.qr(X)
    assert np.allclose(Q.T @ Q, np.eye(8), atol=1e-10)
    # smoothness: neighbouring vertices move together far more than random fields do
    field = Q[:, 0].reshape(-1, 3)
    edge_diff = np.linalg.norm(field[f[:, 0]] - field[f[:, 1]], axis=1).mean()
    assert edge_diff < 0.3 * np.linalg.norm(field, axis=1).mean()


def test_prior_statistics():
    p = fit.sample_params(np.random.Generator(np.random.PCG64(1)), 4000, 80, torch.device("cpu"))
    assert abs(p.alpha.std().item() - 0.5) < 0.01
    assert (p.sh[:, [0, 9, 18]] >= 0.6).all() and (p.sh[:, [0, 9, 18]] <= 1.0).all()
    assert abs(p.pose[:, :3].std().item() - 0.1) < 0.005
    x = p.flat()
    q = fit.Params.from_flat(x, 80)
    assert torch.equal(q.alpha, p.alpha) and torch.equal(q.sh, p.sh) and torch.equal(q.pose, p.pose)
