"""The CG entry point that allocates nothing
(sil_minimize_strong_formulation_residual_noalloc) against the allocating
one, through the C ABI: the same kernels in the same order, so the same
iterations and bit-identical solutions; two calls in a row, the second on a
smaller problem; sizes beyond the scratch array refused."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def settings(L, precond, tol=1e-10, max_iterations=20000):
    st = L.SolverSettings()
    L.lib().sil_default_solver_settings(ctypes.byref(st))
    st.max_iterations = max_iterations
    st.relative_tolerance = tol
    st.precision = L.PRECISION_FP64
    st.preconditioner = precond
    return st


def solve_both(L, op, b, st):
    x_ref = torch.full_like(b, float("nan"))
    x_new = torch.full_like(b, float("nan"))
    d_ref, d_new = L.SolverDiagnostics(), L.SolverDiagnostics()
    L.check(L.lib().sil_minimize_strong_formulation_residual(op.handle, L.ptr(b), L.ptr(x_ref), ctypes.byref(st),
                                                             ctypes.byref(d_ref), None))
    L.check(L.lib().sil_minimize_strong_formulation_residual_noalloc(op.handle, L.ptr(b), L.ptr(x_new),
                                                                     ctypes.byref(st), ctypes.byref(d_new), None))
    return x_ref, d_ref.as_dict(), x_new, d_new.as_dict()


def poisson(n, seed):
    from similie_amd import solvers as S

    rng = np.random.default_rng(seed)
    c = np.linspace(0.0, 1.0, n)
    kappa = torch.from_numpy(np.exp(rng.standard_normal((n, n)))).cuda()
    b = np.zeros((n, n))
    b[1:-1, 1:-1] = rng.standard_normal((n - 2, n - 2))
    return S.Poisson2D(c, c, kappa), torch.from_numpy(b.reshape(-1)).cuda()


@pytest.mark.parametrize("precond", ["identity", "jacobi"])
def test_noalloc_matches_the_allocating_solve_on_two_calls(precond):
    from similie_amd import _lib as L

    pc = L.PRECOND_JACOBI if precond == "jacobi" else L.PRECOND_IDENTITY
    st = settings(L, pc)
    for n, seed in ((256, 1), (97, 2)):  # the second call is smaller
        op, b = poisson(n, seed)
        x_ref, d_ref, x_new, d_new = solve_both(L, op, b, st)
        assert d_new["iterations"] == d_ref["iterations"] > 0
        assert d_new["converged"] == d_ref["converged"] == 1 and d_new["verified"] == 1
        assert d_new["true_relative_residual"] == d_ref["true_relative_residual"]
        assert torch.equal(x_new, x_ref), (n, float((x_new - x_ref).abs().max()))


def test_noalloc_magnetostatics_3d():
    from similie_amd import _lib as L
    from similie_amd import solvers as S

    rng = np.random.default_rng(3)
    shape = (24, 20, 28)
    coords = [np.linspace(0.0, 1.0, m) for m in shape]
    cells = (shape[0] - 1) * (shape[1] - 1) * (shape[2] - 1)
    cell_mu = np.where(rng.random(cells) < 0.3, 1000.0, 1.0) * 4e-7 * np.pi
    mu = S.MagnetostaticsOperator3D.nodal_mu(shape, torch.from_numpy(cell_mu).cuda(), 4e-7 * np.pi)
    op = S.MagnetostaticsOperator3D(*coords, mu, gauge_penalty=1e7)
    b = op.rhs(torch.from_numpy(rng.standard_normal((cells, 3))).cuda())
    x_ref, d_ref, x_new, d_new = solve_both(L, op, b, settings(L, L.PRECOND_JACOBI))
    assert d_new["iterations"] == d_ref["iterations"] and torch.equal(x_new, x_ref)
    assert d_new["verified"] == 1


def test_noalloc_refuses_what_it_cannot_hold():
    from similie_amd import _lib as L

    limit = L.lib().sil_noalloc_max_unknowns()
    assert limit == 1 << 19
    n = int(np.ceil(np.sqrt(limit + 1)))
    op, b = poisson(n, 4)
    x = torch.zeros_like(b)
    d = L.SolverDiagnostics()
    st = settings(L, L.PRECOND_JACOBI)
    rc = L.lib().sil_minimize_strong_formulation_residual_noalloc(op.handle, L.ptr(b), L.ptr(x), ctypes.byref(st),
                                                                  ctypes.byref(d), None)
    assert rc == L.SIL_ERR_UNSUPPORTED and b"exceed" in L.lib().sil_last_error()
    st.precision = L.PRECISION_MIXED
    op2, b2 = poisson(33, 5)
    rc = L.lib().sil_minimize_strong_formulation_residual_noalloc(op2.handle, L.ptr(b2), L.ptr(torch.zeros_like(b2)),
                                                                  ctypes.byref(st), ctypes.byref(d), None)
    assert rc == L.SIL_ERR_UNSUPPORTED and b"FP64 only" in L.lib().sil_last_error()
    assert torch.count_nonzero(x) == 0  # nothing ran
