"""Pins of the oracle's 2D magnetostatics (CPU): the table-driven restatement
of MagnetostaticsOperator2D against a hand-derived closed form, its
structure (criterion, symmetry, definiteness), the nonlinear material and
its Jacobian (finite differences, linear-curve reduction), Newton, the
right-hand side and nodal μ on hand-computed cases, and the 2D cell
post-processing on a uniform field and the slice codifferential."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as orc
from test_oracle_postprocess import force_closed_form

MU0 = 4e-7 * np.pi
CURVE = json.loads((Path(__file__).parent / "golden" / "eicore_bh_curve.json").read_text())


def closed_form(shape, mu, x, terms=1.0):
    """ν(r)(A(r) − A(r+e_y)) + ν(r−e_y)(A(r) − A(r−e_y)) + ν(r)(A(r) − A(r+e_x))
    + ν(r−e_x)(A(r) − A(r−e_x)) on interior rows, the neighbour terms only
    where r − e is interior; identity on the boundary"""
    nx, ny = shape
    A = x.reshape(ny, nx)
    nu = 1.0 / mu.reshape(ny, nx)
    out = A.copy()
    for j in range(1, ny - 1):
        for i in range(1, nx - 1):
            v = nu[j, i] * (A[j, i] - A[j + 1, i]) + nu[j, i] * (A[j, i] - A[j, i + 1])
            if j >= 2:
                v += nu[j - 1, i] * (A[j, i] - A[j - 1, i])
            if i >= 2:
                v += nu[j, i - 1] * (A[j, i] - A[j, i - 1])
            out[j, i] = terms * v
    return out.reshape(-1)


def dense(op, n, state=None):
    cols = []
    for c in range(n):
        e = np.zeros(n)
        e[c] = 1.0
        cols.append(op.apply(e, state) if state is not None else op.apply(e))
    return np.array(cols).T


@pytest.mark.parametrize("criterion,terms", [("moments", 1.0), ("potential", 1.0), ("potential_and_moments", 2.0)])
def test_linear_operator_closed_form(criterion, terms):
    rng = np.random.default_rng(0)
    shape = (9, 7)
    mu = rng.uniform(1, 5, 63) * MU0
    x = rng.standard_normal(63)
    op = orc.Magnetostatics2D(shape, mu, criterion=criterion)
    ref = closed_form(shape, mu, x, terms)
    np.testing.assert_allclose(op.apply(x), ref, rtol=1e-13, atol=1e-13 * np.max(np.abs(ref)))


def test_interior_block_is_symmetric_positive_definite():
    rng = np.random.default_rng(1)
    shape = (7, 6)
    n = 42
    mu = rng.uniform(1, 5, n) * MU0
    A = dense(orc.Magnetostatics2D(shape, mu), n)
    inner = [i + 7 * j for j in range(1, 5) for i in range(1, 6)]
    Ai = A[np.ix_(inner, inner)]
    np.testing.assert_allclose(Ai, Ai.T, rtol=1e-14, atol=1e-14 * np.max(np.abs(Ai)))
    assert np.min(np.linalg.eigvalsh(Ai)) > 0
    # boundary rows are the identity
    bnd = [r for r in range(n) if r not in inner]
    np.testing.assert_array_equal(A[bnd], np.eye(n)[bnd])
    # the linear diagonal is the dense one
    np.testing.assert_allclose(orc.Magnetostatics2D(shape, mu).diagonal(), np.diag(A), rtol=1e-15)


def linear_curve(mu_c, top=2.0, n=6):
    b = np.linspace(0.0, top, n)
    return b, b / mu_c


def test_linear_curve_reduces_to_the_linear_operator():
    rng = np.random.default_rng(2)
    shape = (8, 7)
    n = 56
    mu_c = 500 * MU0
    b, h = linear_curve(mu_c)
    op = orc.Magnetostatics2D(shape, np.full(n, MU0), np.ones(n), b, h)
    lin = orc.Magnetostatics2D(shape, np.full(n, mu_c))
    x = rng.standard_normal(n) * 0.2
    np.testing.assert_allclose(op.apply(x), lin.apply(x), rtol=1e-12, atol=1e-12 * np.max(np.abs(lin.apply(x))))
    v = rng.standard_normal(n)
    np.testing.assert_allclose(op.apply(v, x), lin.apply(v), rtol=1e-10, atol=1e-10 * np.max(np.abs(lin.apply(v))))


@pytest.mark.parametrize("level", [0.05, 0.6, 1.5])
def test_jacobian_matches_finite_differences(level):
    rng = np.random.default_rng(3)
    shape = (8, 9)
    n = 72
    ferro = (rng.random(n) < 0.6).astype(float)
    mu = rng.uniform(1, 3, n) * MU0
    op = orc.Magnetostatics2D(shape, mu, ferro, CURVE["b"], CURVE["h"], criterion="potential_and_moments")
    x = rng.standard_normal(n) * level
    v = rng.standard_normal(n)
    eps = 1e-7 * level
    fd = (op.apply(x + eps * v) - op.apply(x - eps * v)) / (2 * eps)
    jv = op.apply(v, x)
    np.testing.assert_allclose(jv, fd, rtol=1e-4, atol=1e-5 * np.max(np.abs(jv)))
    # its diagonal is the dense diagonal
    np.testing.assert_allclose(op.diagonal(x), np.diag(dense(op, n, x)), rtol=1e-13)


def test_newton_converges():
    rng = np.random.default_rng(4)
    shape = (20, 18)
    n = 360
    ferro = np.zeros(n)
    ferro.reshape(18, 20)[6:12, 7:13] = 1.0
    op = orc.Magnetostatics2D(shape, np.full(n, MU0), ferro, CURVE["b"], CURVE["h"])
    jz = np.zeros(19 * 17)
    jz.reshape(17, 19)[7:10, 2:5] = 1e5
    jz.reshape(17, 19)[7:10, 14:17] = -1e5
    b = orc.ms2d_rhs([np.arange(20.0), np.arange(18.0)], jz)
    x, d = op.newton(b, 1e-10, 5000)
    assert d["converged"] and d["newton_steps"] >= 2, d
    r = b - op.apply(x)
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(b) * 1.0001
    # with a linear curve Newton is one exact linear solve
    bl, hl = linear_curve(200 * MU0)
    opl = orc.Magnetostatics2D(shape, np.full(n, MU0), ferro, bl, hl)
    xl, dl = opl.newton(b, 1e-10, 5000)
    assert dl["converged"] and dl["newton_steps"] <= 2


def test_newton_with_spd_isai_of_the_assembled_jacobian():
    """SpdIsai of the Jacobian assembled at every step: the same solution as
    Jacobi-preconditioned Newton, the same step count, fewer inner
    iterations"""
    shape = (20, 18)
    n = 360
    ferro = np.zeros(n)
    ferro.reshape(18, 20)[6:12, 7:13] = 1.0
    op = orc.Magnetostatics2D(shape, np.full(n, MU0), ferro, CURVE["b"], CURVE["h"])
    jz = np.zeros(19 * 17)
    jz.reshape(17, 19)[7:10, 2:5] = 1e5
    jz.reshape(17, 19)[7:10, 14:17] = -1e5
    b = orc.ms2d_rhs([np.arange(20.0), np.arange(18.0)], jz)
    xj, dj = op.newton(b, 1e-10, 5000)
    xi, di = op.newton(b, 1e-10, 5000, jacobi="spd_isai")
    assert dj["converged"] and di["converged"], (dj, di)
    assert abs(di["newton_steps"] - dj["newton_steps"]) <= 1
    assert di["iterations"] < dj["iterations"], (di, dj)
    assert np.max(np.abs(xi - xj)) <= 1e-7 * np.max(np.abs(xj))
    r = b - op.apply(xi)
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(b) * 1.0001


def test_rhs_and_nodal_mu_hand_values():
    coords = [np.array([0.0, 1.0, 3.0]), np.array([0.0, 2.0, 3.0])]
    # cell areas (x fastest): 1·2, 2·2, 1·1, 2·1
    jz = np.array([1.0, 2.0, 3.0, 4.0])
    rhs = orc.ms2d_rhs(coords, jz)
    expect = np.zeros(9)
    expect[4] = 0.25 * (2 * 1 + 4 * 2 + 1 * 3 + 2 * 4)
    np.testing.assert_array_equal(rhs, expect)
    assert not np.any(orc.ms2d_rhs(coords, jz, criterion="potential"))
    mu = orc.ms2d_nodal_mu((3, 3), np.array([1.0, 2.0, 3.0, 4.0]), 7.0)
    np.testing.assert_array_equal(mu, [1.0, 1.5, 2.0, 2.0, 2.5, 3.0, 3.0, 3.5, 4.0])


def test_post_process_uniform_field_and_force():
    rng = np.random.default_rng(5)
    coords = [np.cumsum(np.concatenate([[0.3], rng.uniform(0.5, 1.5, 8)])),
              np.cumsum(np.concatenate([[-0.2], rng.uniform(0.5, 1.5, 6)]))]
    Y, X = np.meshgrid(coords[1], coords[0], indexing="ij")
    B0 = (0.7, -1.3)
    az = (B0[0] * Y - B0[1] * X).reshape(-1)  # B_x = ∂_y A_z, B_y = −∂_x A_z
    cells = 8 * 6
    mu = rng.uniform(1, 3, cells) * MU0
    B, H, T, F = orc.ms2d_post_process(coords, az, mu)
    np.testing.assert_allclose(B, np.broadcast_to([B0[0], B0[1], 0.0], B.shape), rtol=0, atol=1e-12 * 1.3)
    np.testing.assert_allclose(H, B / mu[:, None], rtol=1e-15)
    # random potential with nonlinear cells: the force is −δ of the stress
    az = rng.standard_normal(az.size)
    nl = (rng.random(cells) < 0.5).astype(float)
    B, H, T, F = orc.ms2d_post_process(coords, az, mu, nl, CURVE["b"], CURVE["h"])
    ref = force_closed_form(coords + [np.array([0.0, 1.0])], T)
    np.testing.assert_allclose(F, ref, rtol=0, atol=1e-12 * np.max(np.abs(ref)))
    assert not np.any(B[:, 2]) and not np.any(T[:, 4]) and not np.any(T[:, 5])


def csr_dense(rp, cols, vals, n):
    d = np.zeros((n, n))
    for r in range(n):
        d[r, cols[rp[r]:rp[r + 1]]] += vals[rp[r]:rp[r + 1]]
    return d


@pytest.mark.parametrize("criterion", ["moments", "potential", "potential_and_moments"])
@pytest.mark.parametrize("nonlinear", [False, True])
def test_assembled_matrix_is_the_operator(criterion, nonlinear):
    """The 2D assembly (the reference's assemble_matrix_data restated) is the
    matrix-free operator — or its Jacobian at a state — column by column;
    columns ascend, no stored zeros, at most 7 per row; the linear matrix is
    symmetric."""
    rng = np.random.default_rng(17)
    shape = (8, 9)
    n = 72
    mu = rng.uniform(1, 3, n) * MU0
    if nonlinear:
        ferro = (rng.random(n) < 0.6).astype(float)
        op = orc.Magnetostatics2D(shape, mu, ferro, CURVE["b"], CURVE["h"], criterion=criterion)
        state = rng.standard_normal(n)
    else:
        op = orc.Magnetostatics2D(shape, mu, criterion=criterion)
        state = None
    rp, cols, vals = op.csr(state)
    A = csr_dense(rp, cols, vals, n)
    D = dense(op, n, state)
    np.testing.assert_allclose(A, D, rtol=0, atol=1e-13 * np.max(np.abs(D)))
    assert all(np.all(np.diff(cols[rp[r]:rp[r + 1]]) > 0) for r in range(n))
    assert np.all(vals != 0.0) and np.max(np.diff(rp)) <= 7
    if not nonlinear:  # the interior block (boundary rows are the identity)
        inner = [i + 8 * j for j in range(1, 8) for i in range(1, 7)]
        Ai = A[np.ix_(inner, inner)]
        np.testing.assert_allclose(Ai, Ai.T, rtol=0, atol=1e-14 * np.max(np.abs(Ai)))
