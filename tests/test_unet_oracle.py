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


def test_split_bf16_is_a_16_bit_two_term_sum():
    """split_bf16 holds a value to 16 significant bits: exact on values that
    have at most 16 (16-bit integers times powers of two), within
    2^-16 |a| on all, and both terms are BF16 values; rounding to BF16 alone
    (a dropped low term) misses the bound by far."""
    rng = np.random.default_rng(7)
    a = rng.standard_normal(200000) * np.exp(rng.uniform(-20, 20, 200000))
    s = U.split_bf16(a)
    a32 = U.f32(a)
    assert np.all(np.abs(a32 - s) <= 2.0 ** -16 * np.abs(a32))
    hi = U.bf16(a32)
    assert np.array_equal(U.bf16(s - hi), s - hi)  # the low term is a BF16 value
    assert np.max(np.abs(a32 - hi) / np.abs(a32)) > 2.0 ** -10
    # 16-bit integers times powers of two: 16 significant bits exactly
    k = rng.integers(2 ** 15, 2 ** 16, 1000).astype(np.float64)
    exact = k * np.exp2(rng.integers(-40, 20, 1000)) * rng.choice([-1.0, 1.0], 1000)
    assert np.array_equal(U.split_bf16(exact), exact)
    assert not np.array_equal(U.bf16(exact), exact)
    assert np.array_equal(U.split_bf16(np.array([0.0, 1.0, -2.5, 3.0])), [0.0, 1.0, -2.5, 3.0])


def test_input_layer_split_form_against_fp64():
    """conv_in_split32 (the input layer as the GPU forms it) against the FP64
    convolution: within the bound the 16-bit operands and the FP32 chain
    allow, 2^-15 + 20·2^-24 of Σ|x·w| per output; the same layer on BF16
    operands (no low terms) lies far outside it."""
    rng = np.random.default_rng(8)
    x = rng.standard_normal((2, 12, 10, 2)).astype(np.float32)
    w = (rng.standard_normal((9, U.C, 2)) * 0.3).astype(np.float32)
    b = (rng.standard_normal(U.C) * 0.1).astype(np.float32)
    ref = U.conv(x.astype(np.float64), w.astype(np.float64), b.astype(np.float64))
    mag = U.conv(np.abs(x).astype(np.float64), np.abs(w).astype(np.float64), np.abs(b).astype(np.float64))
    bound = (2.0 ** -15 + 20 * 2.0 ** -24) * mag
    got = U.conv_in_split32(x, w, b)
    assert np.all(np.abs(got - ref) <= bound)
    assert np.max(np.abs(got - ref) / bound) > 1e-3  # it is an approximation, not the FP64 sum
    coarse = U.conv(U.bf16(x), U.bf16(w), U.f32(b))
    assert np.max(np.abs(coarse - ref) / bound) > 50


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


def test_planted_mistakes_exceed_the_end_to_end_gradient_bound():
    """The end-to-end gradient test on the GPU (test_gpu_learned.py) holds
    every parameter block within GRAD_TOL_EMULATED of the emulating oracle,
    a bound set by how far BF16 rounding alone moves a 12-layer network.  On
    that test's input, mistakes planted in the oracle's parameters — a
    swapped tap pair, transposed or flipped weights, a dropped bias, swapped
    input channels, one input channel of the last 32→32 layer dropped — move
    some block by more than twice the bound."""
    from test_gpu_learned import GRAD_TOL_EMULATED, make_params, rel

    B, H, W = 2, 32, 32
    rng = np.random.default_rng(21)
    p64 = make_params(rng).astype(np.float64)
    x = rng.standard_normal((B, H, W, 2)).astype(np.float32)
    dout = rng.standard_normal((B, H, W)).astype(np.float32)
    blocks = [slice(0, U.N_IN)] + [slice(U.layer_slices(l)[0].start, U.layer_slices(l)[1].stop)
                                    for l in range(1, U.N_LAYERS + 1)] + [slice(U.OUT_BASE, U.N_PARAMS)]

    def grad(p):
        _, act = U.forward(p, x, emulate=True)
        return U.backward(p, act, dout, emulate=True)

    def layer(p, l, f):
        ws, _ = U.layer_slices(l)
        p[ws] = f(p[ws].reshape(9, U.C, U.C).copy()).ravel()

    def swap_taps(w):
        w[[7, 8]] = w[[8, 7]]
        return w

    def drop_channel(w):
        w[:, :, U.C - 1] = 0
        return w

    mistakes = {
        "layer 5 taps 7 and 8 swapped": lambda p: layer(p, 5, swap_taps),
        "layer 9 weights transposed": lambda p: layer(p, 9, lambda w: w.transpose(0, 2, 1)),
        "layer 7 kernel flipped": lambda p: layer(p, 7, lambda w: w[::-1]),
        "layer 3 bias dropped": lambda p: p.__setitem__(U.layer_slices(3)[1], 0.0),
        "input channels swapped": lambda p: p.__setitem__(
            slice(0, 9 * U.C * 2), p[: 9 * U.C * 2].reshape(9, U.C, 2)[..., ::-1].ravel()),
        "layer 12 input channel dropped": lambda p: layer(p, 12, drop_channel),
    }
    g = grad(p64)
    for name, plant in mistakes.items():
        p = p64.copy()
        plant(p)
        gm = grad(p)
        worst = max(rel(gm[s], g[s]) for s in blocks)
        assert worst > 2 * GRAD_TOL_EMULATED, (name, worst)
