"""Pins the oracle's magnetostatics operator, its assembly rules and CG.

No reference test pins these numbers (the reference checks only the
inductance of a gmsh model within 10 %), so the checks are properties the
mathematics fixes: the moment stencils equal the coboundary of the 1-form,
the operator is symmetric and positive definite, gradients see only the
gauge term, boundary rows are the identity, constant fields assemble to
closed forms, and CG reproduces a dense direct solve.
"""
import numpy as np
import pytest

import oracle_lib as orc


def grid_coords(shape, rng=None, stretch=False):
    out = []
    for n in shape:
        if stretch:
            h = rng.uniform(0.5, 1.5, n - 1)
            out.append(np.concatenate([[0.0], np.cumsum(h)]) / 10.0)
        else:
            out.append(np.linspace(0.0, 1.0, n))
    return out


def test_moment_stencils_are_the_coboundary_of_the_potential():
    """B = (dA)_yz, −(dA)_xz, (dA)_xy (magnetostatics_quantities.hpp:173-196):
    with the gauge term off, unit permeability and unit induction Hodge
    factor, ⟨x, A y⟩ = Σ_interior nodes B(x)·B(y) over interior potentials."""
    rng = np.random.default_rng(1)
    shape = (6, 5, 7)
    coords = [np.arange(n, dtype=float) for n in shape]  # unit spacing: w_B = 1
    mu = np.ones(int(np.prod(shape)))
    nx, ny, nz = shape
    x = rng.standard_normal((nz, ny, nx, 3))
    y = rng.standard_normal((nz, ny, nx, 3))
    for a in (x, y):
        a[0], a[-1], a[:, 0], a[:, -1], a[:, :, 0], a[:, :, -1] = 0, 0, 0, 0, 0, 0
    Ay = orc.ms_apply(coords, mu, 0.0, y.ravel())

    def induction(a):
        # DOF order 3(i + nx(j + ny k)) + c; reorder to the oracle's [x][y][z][c]
        arr = np.transpose(a, (2, 1, 0, 3))
        d = orc.coboundary(3, 1, list(shape), arr, list(shape))  # XY, XZ, YZ
        B = np.stack([d[..., 2], -d[..., 1], d[..., 0]], axis=-1)
        return B[1:-1, 1:-1, 1:-1]

    lhs = np.dot(x.ravel(), Ay)
    rhs = np.sum(induction(x) * induction(y))
    assert np.isclose(lhs, rhs, rtol=1e-12)


def interior_mask(shape):
    nx, ny, nz = shape
    m = np.zeros((nz, ny, nx, 3), dtype=bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m.ravel()


@pytest.mark.parametrize("stretch", [False, True])
def test_operator_symmetric_positive_definite_with_identity_boundary(stretch):
    rng = np.random.default_rng(2)
    shape = (5, 6, 5)
    coords = grid_coords(shape, rng, stretch)
    mu = rng.uniform(1.0, 1000.0, int(np.prod(shape)))
    n = 3 * int(np.prod(shape))
    M = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        M[:, j] = orc.ms_apply(coords, mu, 20.0, e)
    inner = interior_mask(shape)
    A = M[np.ix_(inner, inner)]
    np.testing.assert_allclose(A, A.T, rtol=1e-12, atol=1e-12 * np.abs(A).max())
    assert np.linalg.eigvalsh(A).min() > 0
    bnd = ~inner
    np.testing.assert_array_equal(M[np.ix_(bnd, bnd)], np.eye(bnd.sum()))
    assert np.all(M[np.ix_(inner, bnd)] == 0) and np.all(M[np.ix_(bnd, inner)] == 0)
    d = orc.ms_diagonal(coords, mu, 20.0)
    np.testing.assert_allclose(d, np.diag(M), rtol=1e-14)


def test_gradients_only_see_the_gauge_term():
    """curl∘grad = 0: a discrete gradient supported in the interior is mapped
    to P·w_A·x alone."""
    rng = np.random.default_rng(3)
    shape = (8, 7, 9)
    coords = grid_coords(shape, rng, True)
    nx, ny, nz = shape
    mu = rng.uniform(1.0, 10.0, nx * ny * nz)
    phi = np.zeros((nx, ny, nz, 1))
    phi[2:-2, 2:-2, 2:-2, 0] = rng.standard_normal((nx - 4, ny - 4, nz - 4))
    grad = orc.coboundary(3, 0, list(shape), phi, list(shape))  # [x][y][z][3]
    x = np.transpose(grad, (2, 1, 0, 3)).ravel()
    P = 7.5
    out = orc.ms_apply(coords, mu, P, x)

    def ls(c, i):
        n = len(c)
        if i == 0:
            return c[1] - c[0]
        if i == n - 1:
            return c[n - 1] - c[n - 2]
        return 0.5 * (c[i + 1] - c[i - 1])

    expect = np.zeros((nz, ny, nx, 3))
    X = x.reshape(nz, ny, nx, 3)
    for k in range(1, nz - 1):
        for j in range(1, ny - 1):
            for i in range(1, nx - 1):
                dx, dy, dz = ls(coords[0], i), ls(coords[1], j), ls(coords[2], k)
                w = (dy * dz / dx, dx * dz / dy, dx * dy / dz)
                for c in range(3):
                    expect[k, j, i, c] = P * w[c] * X[k, j, i, c]
    got = out.reshape(nz, ny, nx, 3)
    inner = (slice(1, -1),) * 3
    np.testing.assert_allclose(got[inner], expect[inner], rtol=1e-10, atol=1e-9 * np.abs(expect).max())


def test_nodal_mu_and_rhs_closed_forms():
    rng = np.random.default_rng(4)
    shape = (5, 4, 6)
    nx, ny, nz = shape
    cells = (nx - 1) * (ny - 1) * (nz - 1)
    mu = orc.ms_nodal_mu(shape, np.full(cells, 3.0), 1.0)
    np.testing.assert_allclose(mu, 3.0)
    cm = np.arange(cells, dtype=float)
    mu = orc.ms_nodal_mu(shape, cm, 1.0).reshape(nz, ny, nx)
    cmr = cm.reshape(nz - 1, ny - 1, nx - 1)
    assert np.isclose(mu[1, 1, 1], cmr[0:2, 0:2, 0:2].mean())
    assert np.isclose(mu[0, 0, 0], cmr[0, 0, 0])
    coords = grid_coords(shape, rng, True)
    J = np.tile([1.0, 2.0, 3.0], (cells, 1))
    b = orc.ms_rhs(coords, J).reshape(nz, ny, nx, 3)

    def dual(c, i):
        return 0.5 * (c[i + 1] - c[i - 1])

    i, j, k = 2, 1, 3
    dx, dy, dz = dual(coords[0], i), dual(coords[1], j), dual(coords[2], k)
    np.testing.assert_allclose(b[k, j, i], [1.0 * dy * dz, 2.0 * dx * dz, 3.0 * dx * dy], rtol=1e-13)
    assert np.all(b[0] == 0) and np.all(b[:, :, -1] == 0)


def test_cg_matches_dense_solve_on_poisson():
    rng = np.random.default_rng(5)
    nx, ny = 9, 8
    hx = rng.uniform(0.5, 1.5, nx)
    hy = rng.uniform(0.5, 1.5, ny)
    kappa = np.exp(rng.standard_normal((nx, ny)))
    b = np.zeros((nx, ny))
    b[1:-1, 1:-1] = rng.standard_normal((nx - 2, ny - 2))
    x, it, rn, bn = orc.poisson2d_cg(hx, hy, kappa, b, 1e-13, 500)
    inner = [(i, j) for i in range(1, nx - 1) for j in range(1, ny - 1)]
    M = np.zeros((len(inner), len(inner)))
    for col, (i, j) in enumerate(inner):
        e = np.zeros((nx, ny))
        e[i, j] = 1.0
        o = orc.poisson2d_apply(hx, hy, kappa, e)
        M[:, col] = [o[p, q] for p, q in inner]
    np.testing.assert_allclose(M, M.T, rtol=1e-13)
    xd = np.linalg.solve(M, np.array([b[p, q] for p, q in inner]))
    np.testing.assert_allclose([x[p, q] for p, q in inner], xd, rtol=1e-9, atol=1e-11)
    assert rn <= 1e-13 * bn and it <= 3 * len(inner)


def test_cg_solves_small_system_within_n_iterations():
    """n ≤ 10 unknowns: exact solution in at most n iterations."""
    nx, ny = 5, 4  # 3 × 2 = 6 interior unknowns
    hx, hy = np.ones(nx), np.ones(ny)
    b = np.zeros((nx, ny))
    b[1:-1, 1:-1] = np.arange(1.0, 7.0).reshape(3, 2)
    x, it, rn, bn = orc.poisson2d_cg(hx, hy, None, b, 1e-14, 50)
    assert it <= 6
    r = b - orc.poisson2d_apply(hx, hy, None, x)
    assert np.linalg.norm(r) <= 1e-13 * np.linalg.norm(b)


def test_magnetostatics_cg_converges_to_verified_residual():
    rng = np.random.default_rng(6)
    shape = (7, 6, 8)
    coords = grid_coords(shape, rng, False)
    nx, ny, nz = shape
    mu = np.where(rng.random(nx * ny * nz) < 0.3, 1000.0, 1.0)
    cells = (nx - 1) * (ny - 1) * (nz - 1)
    J = rng.standard_normal((cells, 3))
    b = orc.ms_rhs(coords, J)
    for diag in (None, orc.ms_diagonal(coords, mu, 20.0)):
        x, it, rn, bn = orc.ms_cg(coords, mu, 20.0, b, 1e-10, 5000, diag=diag)
        r = b - orc.ms_apply(coords, mu, 20.0, x)
        assert np.linalg.norm(r) <= 2e-10 * np.linalg.norm(b)
