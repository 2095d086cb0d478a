"""Pins the oracle's assembled magnetostatics matrix and its SPD sparse
approximate inverse (CPU only).

The reference holds no golden matrix for either, so the checks are what the
mathematics fixes: the assembled matrix equals the matrix-free operator
column by column (brute force on unit vectors), it is symmetric, its rows
carry the 13-entry curl-curl pattern derived by hand, a uniform grid gives
closed-form entries; the factorized approximate inverse is exact when its
pattern is the full lower triangle, reduces to 1/√a_ii on a diagonal
matrix, and satisfies diag(G A Gᵀ) = 1 and (G A)[i, J_i∖{i}] = 0 — the two
defining conditions of FSAI — on the magnetostatics matrix and on a 2D
Laplacian.  A dropped scaling, a transposed system or a wrong triangle
fails one of these.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle_lib as orc


def stretched(shape, rng):
    return [np.concatenate([[0.0], np.cumsum(rng.uniform(0.5, 1.5, n - 1))]) / 10.0 for n in shape]


def csr(rp, cols, vals, n):
    return sp.csr_matrix((vals, cols, rp), shape=(n, n))


def dense_from_apply(coords, mu, penalty, y_sign=1.0):
    n = 3 * int(np.prod([len(c) for c in coords]))
    A = np.zeros((n, n))
    e = np.zeros(n)
    for j in range(n):
        e[j] = 1.0
        A[:, j] = orc.ms_apply(coords, mu, penalty, e, y_sign)
        e[j] = 0.0
    return A


@pytest.mark.parametrize("y_sign", [1.0, -1.0])
def test_assembled_matrix_equals_the_matrix_free_operator(y_sign):
    rng = np.random.default_rng(3)
    shape = (5, 4, 6)
    coords = stretched(shape, rng)
    mu = rng.uniform(0.5, 3.0, int(np.prod(shape)))
    rp, cols, vals = orc.ms_assemble_csr(coords, mu, 7.5, y_sign)
    n = len(rp) - 1
    A = csr(rp, cols, vals, n).toarray()
    D = dense_from_apply(coords, mu, 7.5, y_sign)
    scale = np.abs(D).max()
    assert np.abs(A - D).max() <= 1e-14 * scale
    # every nonzero of the operator is stored; a stored entry the operator
    # lacks can only be an exact cancellation of nonzero contributions
    stored = np.zeros((n, n), dtype=bool)
    stored[np.repeat(np.arange(n), np.diff(rp)), cols] = True
    assert not np.any((np.abs(D) > 1e-14 * scale) & ~stored)


def test_assembled_rows_ascending_symmetric_and_boundary_identity():
    rng = np.random.default_rng(4)
    shape = (6, 5, 7)
    coords = stretched(shape, rng)
    mu = rng.uniform(0.5, 3.0, int(np.prod(shape)))
    rp, cols, vals = orc.ms_assemble_csr(coords, mu, 20.0)
    n = len(rp) - 1
    for r in range(n):
        c = cols[rp[r]:rp[r + 1]]
        assert np.all(np.diff(c) > 0)
    A = csr(rp, cols, vals, n)
    assert abs(A - A.T).max() <= 1e-15 * abs(A).max()
    nx, ny, nz = shape
    for r in range(n):
        node = r // 3
        i, j, k = node % nx, (node // nx) % ny, node // (nx * ny)
        if i in (0, nx - 1) or j in (0, ny - 1) or k in (0, nz - 1):
            assert rp[r + 1] - rp[r] == 1 and cols[rp[r]] == r and vals[rp[r]] == 1.0


def test_assembled_row_pattern_is_the_curl_curl_stencil():
    """An A_x row two nodes from every face: A_x at n, n±e_y, n±e_z; A_y at
    n, n+e_x, n−e_y, n+e_x−e_y; A_z at n, n+e_x, n−e_z, n+e_x−e_z (13), and
    the cyclic patterns for A_y and A_z."""
    shape = (7, 7, 7)
    coords = [np.linspace(0.0, 1.0, s) for s in shape]
    mu = np.ones(int(np.prod(shape)))
    rp, cols, vals = orc.ms_assemble_csr(coords, mu, 1.0)
    nx, ny, _ = shape

    def dof(i, j, k, c):
        return 3 * (i + nx * (j + ny * k)) + c

    i, j, k = 3, 3, 3
    expect = {
        0: [(0, 0, 0, 0), (0, 1, 0, 0), (0, -1, 0, 0), (0, 0, 1, 0), (0, 0, -1, 0),
            (0, 0, 0, 1), (1, 0, 0, 1), (0, -1, 0, 1), (1, -1, 0, 1),
            (0, 0, 0, 2), (1, 0, 0, 2), (0, 0, -1, 2), (1, 0, -1, 2)],
        1: [(0, 0, 0, 1), (1, 0, 0, 1), (-1, 0, 0, 1), (0, 0, 1, 1), (0, 0, -1, 1),
            (0, 0, 0, 0), (0, 1, 0, 0), (-1, 0, 0, 0), (-1, 1, 0, 0),
            (0, 0, 0, 2), (0, 1, 0, 2), (0, 0, -1, 2), (0, 1, -1, 2)],
        2: [(0, 0, 0, 2), (1, 0, 0, 2), (-1, 0, 0, 2), (0, 1, 0, 2), (0, -1, 0, 2),
            (0, 0, 0, 0), (0, 0, 1, 0), (-1, 0, 0, 0), (-1, 0, 1, 0),
            (0, 0, 0, 1), (0, 0, 1, 1), (0, -1, 0, 1), (0, -1, 1, 1)],
    }
    for c, taps in expect.items():
        r = dof(i, j, k, c)
        got = sorted(cols[rp[r]:rp[r + 1]].tolist())
        assert got == sorted(dof(i + a, j + b, k + d, cc) for a, b, d, cc in taps)


def test_uniform_grid_closed_form_entries():
    """Spacing h, permeability μ, penalty P: (A_x, A_x) diagonal = P·h +
    4/(μ h) (two moments, each through two sampled nodes), (A_x, A_x at
    n+e_y) = −1/(μ h), and through B_z = A_y(s+e_x) − A_y(s) − A_x(s+e_y) +
    A_x(s) at s = n: (A_x, A_y at n) = −1/(μ h), (A_x, A_y at n+e_x) =
    +1/(μ h)."""
    shape = (7, 7, 7)
    h = 0.25
    coords = [np.arange(s) * h for s in shape]
    mu_v, P = 2.0, 3.0
    mu = np.full(int(np.prod(shape)), mu_v)
    rp, cols, vals = orc.ms_assemble_csr(coords, mu, P)
    nx, ny, _ = shape
    r = 3 * (3 + nx * (3 + ny * 3))
    row = dict(zip(cols[rp[r]:rp[r + 1]].tolist(), vals[rp[r]:rp[r + 1]].tolist()))
    assert row[r] == pytest.approx(P * h + 4.0 / (mu_v * h), rel=1e-15)
    assert row[r + 3 * nx] == pytest.approx(-1.0 / (mu_v * h), rel=1e-15)
    assert row[r + 1] == pytest.approx(-1.0 / (mu_v * h), rel=1e-15)
    assert row[r + 3 + 1] == pytest.approx(1.0 / (mu_v * h), rel=1e-15)


def test_assembled_diagonal_matches_probed_diagonal():
    rng = np.random.default_rng(5)
    shape = (6, 5, 7)
    coords = stretched(shape, rng)
    mu = rng.uniform(0.5, 3.0, int(np.prod(shape)))
    rp, cols, vals = orc.ms_assemble_csr(coords, mu, 20.0)
    A = csr(rp, cols, vals, len(rp) - 1)
    d = orc.ms_diagonal(coords, mu, 20.0)
    np.testing.assert_allclose(A.diagonal(), d, rtol=1e-14, atol=0)


# ---------------------------------------------------------------- FSAI ---

def random_spd(n, rng):
    B = rng.standard_normal((n, n))
    return B @ B.T + n * np.eye(n)


def test_spd_isai_full_pattern_is_the_exact_inverse():
    rng = np.random.default_rng(6)
    n = 12
    A = random_spd(n, rng)
    M = sp.csr_matrix(A)
    M.sort_indices()
    grp, gc, gv = orc.spd_isai(M.indptr, M.indices, M.data)
    G = csr(grp, gc, gv, n).toarray()
    assert np.allclose(np.triu(G, 1), 0.0)
    inv = np.linalg.inv(A)
    np.testing.assert_allclose(G.T @ G, inv, rtol=0, atol=1e-12 * np.abs(inv).max())
    # G is the inverse of the Cholesky factor: G A Gᵀ = I
    np.testing.assert_allclose(G @ A @ G.T, np.eye(n), atol=1e-12)


def test_spd_isai_diagonal_matrix():
    a = np.array([4.0, 9.0, 0.25, 2.0])
    M = sp.csr_matrix(np.diag(a))
    grp, gc, gv = orc.spd_isai(M.indptr, M.indices, M.data)
    np.testing.assert_array_equal(gc, np.arange(4))
    np.testing.assert_allclose(gv, 1.0 / np.sqrt(a), rtol=1e-15)


def fsai_conditions(A, G):
    n = A.shape[0]
    GA = (G @ A).tocsr()
    GAGt = (GA @ G.T).tocsr()
    np.testing.assert_allclose(GAGt.diagonal(), np.ones(n), rtol=0, atol=1e-13)
    scale = abs(GA).max()
    for i in range(n):
        Jc = G.indices[G.indptr[i]:G.indptr[i + 1]]
        lower = A.indices[A.indptr[i]:A.indptr[i + 1]]
        np.testing.assert_array_equal(Jc, lower[lower <= i])
        row = GA.getrow(i).toarray().ravel()
        off = Jc[Jc != i]
        assert np.abs(row[off]).max(initial=0.0) <= 1e-13 * scale


def test_spd_isai_conditions_on_the_magnetostatics_matrix():
    rng = np.random.default_rng(7)
    shape = (6, 5, 7)
    coords = stretched(shape, rng)
    mu = rng.uniform(0.5, 3.0, int(np.prod(shape)))
    rp, cols, vals = orc.ms_assemble_csr(coords, mu, 20.0)
    n = len(rp) - 1
    A = csr(rp, cols, vals, n)
    grp, gc, gv = orc.spd_isai(rp, cols, vals)
    fsai_conditions(A, csr(grp, gc, gv, n))


def test_spd_isai_conditions_on_a_2d_laplacian():
    m = 9
    T = sp.diags([-1.0, 2.5, -1.0], [-1, 0, 1], shape=(m, m))
    A = (sp.kron(sp.eye(m), T) + sp.kron(T, sp.eye(m))).tocsr()
    A.sort_indices()
    grp, gc, gv = orc.spd_isai(A.indptr, A.indices, A.data)
    fsai_conditions(A, csr(grp, gc, gv, A.shape[0]))


def test_spd_isai_rejects_missing_diagonal_and_indefinite_rows():
    A = sp.csr_matrix(np.array([[1.0, 0.0], [1.0, 0.0]]))
    A.eliminate_zeros()
    with pytest.raises(ValueError, match="-2"):
        orc.spd_isai(A.indptr, A.indices, A.data)
    B = sp.csr_matrix(np.array([[1.0, 2.0], [2.0, 1.0]]))
    with pytest.raises(ValueError, match="code 2"):
        orc.spd_isai(B.indptr, B.indices, B.data)
