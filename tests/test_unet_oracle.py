"""Pins the CPU oracle of the learned preconditioner's network: its
gradients against central finite differences, its building blocks against
closed forms, and its loss against the definition it restates."""
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
from oracle import unet_oracle as U  # noqa: E402


def random_params(rng, scale=1.0):
    p = np.zeros(U.N_PARAMS)
    p[: 9 * U.C * 2] = rng.standard_normal(9 * U.C * 2) * np.sqrt(2 / 18) * scale
    p[9 * U.C * 2: U.N_IN] = rng.standard_normal(U.C) * 0.1
    for l in range(1, U.N_LAYERS + 1):
        ws, bs = U.layer_slices(l)
        p[ws] = rng.standard_normal(9 * U.C * U.C) * np.sqrt(2 / (9 * U.C)) * scale
        p[bs] = rng.standard_normal(U.C) * 0.1
    p[U.OUT_BASE:] = rng.standard_normal(U.N_OUT) * np.sqrt(1 / (9 * U.C))
    return p


def test_conv_matches_direct_sum():
    """A single output of conv() against the definition summed by hand."""
    rng = np.random.default_rng(0)
    x = rng.standard_normal((1, 5, 6, 3))
    w = rng.standard_normal((9, 4, 3))
    b = rng.standard_normal(4)
    y = U.conv(x, w, b)
    for (i, j, o) in [(0, 0, 0), (2, 3, 1), (4, 5, 3), (4, 0, 2)]:
        s = b[o]
        for t in range(9):
            dy, dx = divmod(t, 3)
            ii, jj = i + dy - 1, j + dx - 1
            if 0 <= ii < 5 and 0 <= jj < 6:
                s += w[t, o] @ x[0, ii, jj]
        assert y[0, i, j, o] == pytest.approx(s, rel=1e-13, abs=1e-13)


def test_conv_backward_is_the_adjoint():
    """<dz, conv(x)> = <conv_bwd dx, x> and the weight gradient is linear in dz."""
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 6, 4, 3))
    w = rng.standard_normal((9, 5, 3))
    dz = rng.standard_normal((2, 6, 4, 5))
    dx, dw, db = U.conv_bwd(x, w, dz)
    lhs = np.sum(dz * U.conv(x, w, np.zeros(5)))
    assert lhs == pytest.approx(np.sum(dx * x), rel=1e-12)
    assert lhs == pytest.approx(np.sum(dw * w), rel=1e-12)
    assert np.allclose(db, dz.sum(axis=(0, 1, 2)))


def test_pool_and_up_are_adjoint_pairs():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((1, 8, 6, 2))
    y = rng.standard_normal((1, 4, 3, 2))
    assert np.sum(U.pool(x) * y) == pytest.approx(np.sum(x * U.pool_bwd(y)), rel=1e-13)
    assert np.sum(U.up(y) * x) == pytest.approx(np.sum(y * U.up_bwd(x)), rel=1e-13)


def test_network_gradient_matches_finite_differences():
    """Central differences of sum(dout * out) on parameters of every layer."""
    rng = np.random.default_rng(3)
    params = random_params(rng)
    x = rng.standard_normal((1, 16, 16, 2))
    dout = rng.standard_normal((1, 16, 16))
    out, act = U.forward(params, x)
    g = U.backward(params, act, dout)
    picks = list(rng.choice(9 * U.C * 2, 3, replace=False)) + [9 * U.C * 2 + 5]
    for l in range(1, U.N_LAYERS + 1):
        ws, bs = U.layer_slices(l)
        picks += list(ws.start + rng.choice(9 * U.C * U.C, 3, replace=False)) + [bs.start + 7]
    picks += [U.OUT_BASE + 11, U.OUT_BASE + 9 * U.C]
    eps = 1e-6
    bad = 0
    for i in picks:
        pp, pm = params.copy(), params.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (np.sum(dout * U.forward(pp, x)[0]) - np.sum(dout * U.forward(pm, x)[0])) / (2 * eps)
        if not np.isclose(g[i], fd, rtol=1e-5, atol=1e-7):
            bad += 1  # a ReLU kink inside [−eps, eps] can flip one difference
    assert bad <= 1, f"{bad} of {len(picks)} finite differences disagree"


def test_direction_loss_definition_and_gradient():
    rng = np.random.default_rng(4)
    B, n = 2, 12
    kappa = np.exp(rng.standard_normal((B, n, n)))
    wx = np.empty_like(kappa)
    wy = np.empty_like(kappa)
    wx[:, :-1, :] = 0.5 * (kappa[:, :-1, :] + kappa[:, 1:, :])
    wx[:, -1, :] = kappa[:, -1, :]
    wy[:, :, :-1] = 0.5 * (kappa[:, :, :-1] + kappa[:, :, 1:])
    wy[:, :, -1] = kappa[:, :, -1]
    e = rng.standard_normal((B, n, n))
    e[:, [0, -1], :] = 0.0
    e[:, :, [0, -1]] = 0.0
    r = U.apply5(e, wx, wy)
    # exact error direction: zero loss
    loss, _ = U.direction_loss(3.0 * e, r, e, wx, wy)
    assert np.allclose(loss, 0.0, atol=1e-12)
    N = rng.standard_normal((B, n, n))
    N[:, [0, -1], :] = 0.0
    N[:, :, [0, -1]] = 0.0
    loss, dN = U.direction_loss(N, r, e, wx, wy)
    # 1 - cos^2 in the A inner product, computed from the definition
    Ae, AN = U.apply5(e, wx, wy), U.apply5(N, wx, wy)
    for b in range(B):
        cos2 = np.sum(e[b] * AN[b]) ** 2 / (np.sum(e[b] * Ae[b]) * np.sum(N[b] * AN[b]))
        assert loss[b] == pytest.approx(1 - cos2, rel=1e-12)
    assert np.all((loss >= 0) & (loss <= 1))
    eps = 1e-6
    for (b, i, j) in [(0, 3, 4), (1, 7, 2), (1, 10, 10)]:
        Np, Nm = N.copy(), N.copy()
        Np[b, i, j] += eps
        Nm[b, i, j] -= eps
        fd = (U.direction_loss(Np, r, e, wx, wy)[0].mean() - U.direction_loss(Nm, r, e, wx, wy)[0].mean()) / (2 * eps)
        assert dN[b, i, j] == pytest.approx(fd, rel=1e-5, abs=1e-9)
