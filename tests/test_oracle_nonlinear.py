"""Pins the oracle's nonlinear magnetostatics (CPU only): the interpolated
B-H curve, the material response of the generated Hamiltonian, the
nonlinear operator, its Jacobian and the Newton solve.

The reference holds no golden output for any of them, and the C++ the
sympy generator writes is not in the tree (it is generated at build time),
so the checks are what the mathematics fixes: the curve interpolates its
samples (the EIcore samples of the reference's inductor example, a
fixture), its derivative functions match finite differences, a linear
curve or a non-ferromagnetic node reduces everything to the linear
operator, the Jacobian matches central differences of the operator and is
symmetric where the Hodge weights are equal, its diagonal is the dense
diagonal, and Newton converges (in one step on a linear curve).
"""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as orc

CURVE = json.loads((Path(__file__).parent / "golden" / "eicore_bh_curve.json").read_text())
BB, HH = np.array(CURVE["b"]), np.array(CURVE["h"])
MU0 = 4e-7 * np.pi


def test_curve_fixture():
    assert len(BB) == len(HH) == 19
    assert BB[0] == 0.0 and HH[0] == 0.0 and np.all(np.diff(BB) > 0) and np.all(np.diff(HH) > 0)


def test_bh_curve_interpolates_its_samples():
    for i in range(1, len(BB)):
        assert orc.bh_eval(BB, HH, "h_from_b", BB[i]) == pytest.approx(HH[i], rel=1e-14)
        assert orc.bh_eval(BB, HH, "h_over_b_from_q", BB[i] ** 2) == pytest.approx(HH[i] / BB[i], rel=1e-14)
        mid = 0.5 * (BB[i - 1] + BB[i])
        assert orc.bh_eval(BB, HH, "h_from_b", mid) == pytest.approx(0.5 * (HH[i - 1] + HH[i]), rel=1e-13)
        assert orc.bh_eval(BB, HH, "dh_db_from_b", mid) == pytest.approx(
            (HH[i] - HH[i - 1]) / (BB[i] - BB[i - 1]), rel=1e-13)
    # beyond the last sample: the last segment's slope; below the first: h0
    slope = (HH[-1] - HH[-2]) / (BB[-1] - BB[-2])
    assert orc.bh_eval(BB, HH, "h_from_b", 2.5) == pytest.approx(HH[-2] + slope * (2.5 - BB[-2]), rel=1e-14)
    assert orc.bh_eval(BB, HH, "h_from_b", -1.0) == HH[0]
    # ν₀ = ν₁ and the zero-field limit of h/b is the first slope
    assert orc.bh_eval(BB, HH, "nu", 0) == orc.bh_eval(BB, HH, "nu", 1) == pytest.approx(HH[1] / BB[1], rel=1e-15)
    assert orc.bh_eval(BB, HH, "h_over_b_from_q", 0.0) == pytest.approx(HH[1] / BB[1], rel=1e-15)


@pytest.mark.parametrize("b", [0.1, 0.5, 0.9, 1.3, 1.55, 1.72, 1.78, 2.2])
def test_bh_derivatives_match_finite_differences(b):
    q, eps = b * b, 1e-7 * b * b
    f = lambda x: orc.bh_eval(BB, HH, "h_over_b_from_q", x)  # noqa: E731
    f1 = lambda x: orc.bh_eval(BB, HH, "d_h_over_b_dq_from_q", x)  # noqa: E731
    f2 = orc.bh_eval(BB, HH, "d2_h_over_b_dq2_from_q", q)
    fd1 = (f(q + eps) - f(q - eps)) / (2 * eps)
    fd2 = (f1(q + eps) - f1(q - eps)) / (2 * eps)
    assert f1(q) == pytest.approx(fd1, rel=1e-5, abs=1e-9 * abs(f(q)))
    assert f2 == pytest.approx(fd2, rel=1e-4, abs=1e-9 * abs(f(q)))


def problem(shape, rng, ferro_share=0.5, bh=(BB, HH)):
    coords = [np.concatenate([[0.0], np.cumsum(rng.uniform(0.5, 1.5, n - 1))]) / (n - 1) for n in shape]
    nn = int(np.prod(shape))
    mu = rng.uniform(1.0, 3.0, nn) * MU0
    ferro = (rng.random(nn) < ferro_share).astype(float)
    return orc.Nonlinear(coords, mu, ferro, bh[0], bh[1], 1e5), coords, mu


def state(shape, rng, scale=0.5):
    return scale * rng.standard_normal(3 * int(np.prod(shape)))


def test_linear_curve_reduces_to_the_linear_operator():
    rng = np.random.default_rng(2)
    shape = (6, 5, 7)
    mu_c = 500 * MU0
    lb = np.array([0.0, 0.5, 1.0, 1.5, 2.0, 3.0])
    nl, coords, _ = problem(shape, rng, ferro_share=1.1, bh=(lb, lb / mu_c))
    x, v = state(shape, rng), state(shape, rng)
    mu_lin = np.full(int(np.prod(shape)), mu_c)
    ref = orc.ms_apply(coords, mu_lin, 1e5, x)
    assert np.max(np.abs(nl.apply(x) - ref)) <= 1e-12 * np.max(np.abs(ref))
    refj = orc.ms_apply(coords, mu_lin, 1e5, v)
    assert np.max(np.abs(nl.jacobian_apply(x, v) - refj)) <= 1e-12 * np.max(np.abs(refj))


def test_non_ferromagnetic_nodes_are_linear():
    rng = np.random.default_rng(3)
    shape = (6, 5, 7)
    nl, coords, mu = problem(shape, rng, ferro_share=0.0)
    x, v = state(shape, rng), state(shape, rng)
    ref = orc.ms_apply(coords, mu, 1e5, x)
    assert np.max(np.abs(nl.apply(x) - ref)) <= 1e-14 * np.max(np.abs(ref))
    refj = orc.ms_apply(coords, mu, 1e5, v)
    assert np.max(np.abs(nl.jacobian_apply(x, v) - refj)) <= 1e-14 * np.max(np.abs(refj))


@pytest.mark.parametrize("scale", [0.05, 0.5, 2.0])
def test_jacobian_is_the_derivative_of_the_operator(scale):
    rng = np.random.default_rng(int(scale * 100))
    shape = (7, 6, 8)
    nl, _, _ = problem(shape, rng)
    x, v = state(shape, rng, scale), state(shape, rng)
    jv = nl.jacobian_apply(x, v)
    eps = 1e-7 * scale
    fd = (nl.apply(x + eps * v) - nl.apply(x - eps * v)) / (2 * eps)
    assert np.linalg.norm(fd - jv) <= 1e-6 * np.linalg.norm(jv)


def test_jacobian_symmetric_with_equal_hodge_weights():
    rng = np.random.default_rng(6)
    shape = (6, 6, 6)
    nl, _, _ = problem(shape, rng)
    nl.coords = [np.linspace(0.0, 1.0, n) for n in shape]
    x, u, v = state(shape, rng), state(shape, rng), state(shape, rng)
    a, b = u @ nl.jacobian_apply(x, v), v @ nl.jacobian_apply(x, u)
    assert a == pytest.approx(b, rel=1e-12)


def test_jacobian_diagonal_is_the_dense_diagonal():
    rng = np.random.default_rng(7)
    shape = (4, 5, 4)
    nl, _, _ = problem(shape, rng)
    x = state(shape, rng)
    n = nl.n
    dense = np.zeros(n)
    e = np.zeros(n)
    for j in range(n):
        e[j] = 1.0
        dense[j] = nl.jacobian_apply(x, e)[j]
        e[j] = 0.0
    np.testing.assert_allclose(nl.jacobian_diagonal(x), dense, rtol=1e-15, atol=0)


def rhs_for(nl, rng):
    """a right-hand side F(x*) of a state x* that saturates part of the core"""
    xs = state(tuple(len(c) for c in nl.coords), rng, 1.0)
    return nl.apply(xs)


def test_newton_single_step_on_a_linear_curve():
    rng = np.random.default_rng(8)
    shape = (6, 5, 7)
    mu_c = 800 * MU0
    lb = np.array([0.0, 1.0, 2.0, 4.0])
    nl, _, _ = problem(shape, rng, ferro_share=1.1, bh=(lb, lb / mu_c))
    b = rhs_for(nl, rng)
    x, d = nl.newton(b, 1e-10, 5000)
    assert d["converged"] and d["newton_steps"] <= 2
    assert np.linalg.norm(b - nl.apply(x)) <= 1e-10 * np.linalg.norm(b) * 1.0001


@pytest.mark.parametrize("jacobi", [True, False])
def test_newton_converges_on_the_eicore_curve(jacobi):
    rng = np.random.default_rng(9)
    shape = (7, 6, 8)
    nl, _, _ = problem(shape, rng)
    b = rhs_for(nl, rng)
    x, d = nl.newton(b, 1e-10, 20000, jacobi=jacobi)
    assert d["converged"], d
    assert 2 <= d["newton_steps"] <= 30
    r = np.linalg.norm(b - nl.apply(x))
    assert r <= 1e-10 * d["initial_residual_l2"] * 1.0001
    assert d["final_residual_l2"] == pytest.approx(r, rel=1e-6)


def csr_dense(rp, cols, vals, n):
    d = np.zeros((n, n))
    for r in range(n):
        for q in range(rp[r], rp[r + 1]):
            d[r, cols[q]] += vals[q]
    return d


@pytest.mark.parametrize("scale", [0.05, 1.0])
def test_jacobian_assembly_equals_the_matrix_free_jacobian(scale):
    """The assembled J(state) is the matrix-free J(state) column by column,
    its rows ascend, and the pattern holds at most 23 columns per row (the
    three moments of the four sampled nodes that reach a DOF)."""
    rng = np.random.default_rng(int(10 * scale) + 3)
    shape = (5, 4, 6)
    nl, _, _ = problem(shape, rng)
    x = state(shape, rng, scale)
    rp, cols, vals = nl.jacobian_csr(x)
    n = nl.n
    A = csr_dense(rp, cols, vals, n)
    e = np.zeros(n)
    for j in range(n):
        e[j] = 1.0
        col = nl.jacobian_apply(x, e)
        e[j] = 0.0
        assert np.max(np.abs(A[:, j] - col)) <= 1e-13 * max(np.max(np.abs(col)), 1e-300), j
    assert all(np.all(np.diff(cols[rp[r]:rp[r + 1]]) > 0) for r in range(n))
    assert np.max(np.diff(rp)) <= 23


def test_jacobian_assembly_of_a_linear_curve_is_the_linear_assembly():
    """Without ferromagnetic nodes the Jacobian is the linear operator: the
    assembly reduces to orc_ms_assemble_csr entry for entry."""
    rng = np.random.default_rng(12)
    shape = (6, 5, 7)
    nl, coords, mu = problem(shape, rng)
    nl.ferro = np.zeros_like(nl.ferro)
    x = state(shape, rng)
    rp, cols, vals = nl.jacobian_csr(x)
    rp0, cols0, vals0 = orc.ms_assemble_csr(coords, mu, nl.penalty)
    assert np.array_equal(rp, rp0) and np.array_equal(cols, cols0)
    assert np.array_equal(vals, vals0)


def test_newton_with_spd_isai_of_the_assembled_jacobian():
    """inductor.silpro's setting (SpdIsai with Newton): the Jacobian is
    assembled at every step and its SpdIsai preconditions the inner CG.  On a
    grid of equal spacings in all directions the Jacobian is symmetric
    positive definite; Newton takes the Jacobi path's steps, with fewer inner
    iterations."""
    rng = np.random.default_rng(9)
    shape = (7, 7, 7)
    nl, _, _ = problem(shape, rng)
    nl.coords = [np.linspace(0.0, 1.0, n) for n in shape]
    b = rhs_for(nl, rng)
    xj, dj = nl.newton(b, 1e-10, 20000, jacobi=True)
    xi, di = nl.newton(b, 1e-10, 20000, jacobi="spd_isai")
    assert di["converged"] and dj["converged"]
    assert abs(di["newton_steps"] - dj["newton_steps"]) <= 1
    assert di["iterations"] < dj["iterations"]
    r = np.linalg.norm(b - nl.apply(xi))
    assert r <= 1e-10 * di["initial_residual_l2"] * 1.0001
    rp, cols, vals = nl.jacobian_csr(xi)
    A = csr_dense(rp, cols, vals, nl.n)
    assert np.max(np.abs(A - A.T)) <= 1e-14 * np.max(np.abs(A))
