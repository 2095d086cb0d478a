"""GPU parity of the nonlinear magnetostatics path against the CPU oracle:
the operator F(x), its Jacobian J(x)·v and the Jacobian's diagonal (FP64,
within 1e-12 relative), and the Newton solve (same step count, verified
residual, the oracle's solution).  B-H curve: the EIcore samples of the
reference's inductor example (tests/golden/eicore_bh_curve.json)."""
import json
from pathlib import Path

import numpy as np
import pytest
import torch

import oracle_lib as orc

pytestmark = pytest.mark.gpu

CURVE = json.loads((Path(__file__).parent / "golden" / "eicore_bh_curve.json").read_text())
BB, HH = np.array(CURVE["b"]), np.array(CURVE["h"])
MU0 = 4e-7 * np.pi


@pytest.fixture(scope="module")
def S():
    import similie_amd.solvers as S

    return S


def cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def make(S, shape, seed, bh=(BB, HH), penalty=1e5, ferro_share=0.5):
    rng = np.random.default_rng(seed)
    coords = [np.concatenate([[0.0], np.cumsum(rng.uniform(0.5, 1.5, n - 1))]) / (n - 1) for n in shape]
    nn = int(np.prod(shape))
    mu = rng.uniform(1.0, 3.0, nn) * MU0
    cells = int(np.prod([n - 1 for n in shape]))
    flags = (rng.random(cells) < ferro_share).astype(float)
    frac = S.NonlinearMagnetostaticsOperator3D.nodal_fraction(shape, cuda(flags))
    frac_ref = orc.ms_nodal_mu(shape, flags, 0.0)
    np.testing.assert_allclose(frac.cpu().numpy(), frac_ref, rtol=1e-15)
    op = S.NonlinearMagnetostaticsOperator3D(*coords, cuda(mu), frac, bh[0], bh[1], gauge_penalty=penalty)
    ref = orc.Nonlinear(coords, mu, frac_ref, bh[0], bh[1], penalty)
    return op, ref, rng


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


SHAPES = [(3, 3, 3), (6, 5, 7), (17, 9, 12), (40, 33, 21)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("scale", [0.05, 0.8])
def test_operator_and_jacobian_match_the_oracle(S, shape, scale):
    op, ref, rng = make(S, shape, sum(shape))
    n = op.size()
    x, v = scale * rng.standard_normal(n), rng.standard_normal(n)
    fx = op.apply(cuda(x)).cpu().numpy()
    assert rel(fx, ref.apply(x)) <= 1e-12
    jv = op.apply_jacobian(cuda(x), cuda(v)).cpu().numpy()
    assert rel(jv, ref.jacobian_apply(x, v)) <= 1e-12


@pytest.mark.parametrize("shape", [(6, 5, 7), (17, 9, 12)])
def test_jacobian_diagonal_matches_the_oracle(S, shape):
    op, ref, rng = make(S, shape, 3 + sum(shape))
    x = 0.6 * rng.standard_normal(op.size())
    d = op.jacobian_diagonal(cuda(x)).cpu().numpy()
    np.testing.assert_allclose(d, ref.jacobian_diagonal(x), rtol=1e-12, atol=0)


def test_linear_curve_gives_the_linear_operator(S):
    shape = (17, 9, 12)
    mu_c = 700 * MU0
    lb = np.array([0.0, 0.5, 1.0, 2.0, 4.0])
    op, ref, rng = make(S, shape, 5, bh=(lb, lb / mu_c), ferro_share=1.1)
    rng2 = np.random.default_rng(1)
    coords = [np.array(c) for c in ref.coords]
    lin = S.MagnetostaticsOperator3D(*coords, cuda(np.full(int(np.prod(shape)), mu_c)), gauge_penalty=1e5)
    x = rng2.standard_normal(op.size())
    assert rel(op.apply(cuda(x)).cpu().numpy(), lin.apply(cuda(x)).cpu().numpy()) <= 1e-12


@pytest.mark.parametrize("precond", ["jacobi", "identity"])
def test_newton_matches_the_oracle(S, precond):
    shape = (7, 6, 8)
    op, ref, rng = make(S, shape, 9)
    xs = rng.standard_normal(op.size())
    b = ref.apply(xs)
    tol = 1e-10
    st = S.StrongFormulationSolverSettings(relative_tolerance=tol, max_iterations=20000, preconditioner=precond)
    x, d = S.minimize_strong_formulation_residual(op, cuda(b), settings=st)
    xr, dr = ref.newton(b, tol, 20000, jacobi=precond == "jacobi")
    assert d["converged"] and dr["converged"], (d, dr)
    assert abs(d["outer_iterations"] - dr["newton_steps"]) <= 1, (d, dr)
    assert d["final_relative_residual"] <= tol
    x = x.cpu().numpy()
    r = np.linalg.norm(b - ref.apply(x)) / np.linalg.norm(b)
    assert r <= 1.0001 * tol
    assert rel(x, xr) <= 1e-5


def test_newton_single_step_on_a_linear_curve(S):
    shape = (12, 10, 9)
    mu_c = 900 * MU0
    lb = np.array([0.0, 1.0, 2.0, 4.0])
    op, ref, rng = make(S, shape, 11, bh=(lb, lb / mu_c), ferro_share=1.1)
    b = ref.apply(rng.standard_normal(op.size()))
    x, d = S.minimize_strong_formulation_residual(op, cuda(b), settings=S.StrongFormulationSolverSettings(
        relative_tolerance=1e-10, max_iterations=5000))
    assert d["converged"] and d["outer_iterations"] <= 2


def test_newton_with_mixed_precision_inner_solves(S):
    """FP32 Jacobian applications inside CG, FP64 residual replacement and
    FP64 F(x): the same Newton steps and a verified solution"""
    shape = (17, 9, 12)
    op, ref, rng = make(S, shape, 21)
    b = ref.apply(0.1 * rng.standard_normal(op.size()))
    tol = 1e-10
    res = {}
    for prec in ("fp64", "mixed"):
        st = S.StrongFormulationSolverSettings(relative_tolerance=tol, max_iterations=20000, precision=prec)
        x, d = S.minimize_strong_formulation_residual(op, cuda(b), settings=st)
        assert d["converged"], (prec, d)
        r = np.linalg.norm(b - ref.apply(x.cpu().numpy())) / np.linalg.norm(b)
        assert r <= 1.0001 * tol, (prec, r)
        res[prec] = (x.cpu().numpy(), d)
    assert abs(res["fp64"][1]["outer_iterations"] - res["mixed"][1]["outer_iterations"]) <= 1
    assert rel(res["mixed"][0], res["fp64"][0]) <= 1e-5


def test_newton_stagnation_matches_the_oracle(S):
    """a field strong enough that Newton with backtracking stalls on the
    kinked curve: after the 30 steps the GPU and the oracle stop at the same
    residual"""
    shape = (17, 9, 12)
    op, ref, rng = make(S, shape, 21)
    b = ref.apply(rng.standard_normal(op.size()))
    st = S.StrongFormulationSolverSettings(relative_tolerance=1e-10, max_iterations=20000)
    x, d = S.minimize_strong_formulation_residual(op, cuda(b), settings=st)
    xr, dr = ref.newton(b, 1e-10, 20000)
    assert not d["converged"] and not dr["converged"]
    assert d["outer_iterations"] == dr["newton_steps"] == 30
    assert d["final_residual_l2"] == pytest.approx(dr["final_residual_l2"], rel=1e-6)


def make_cubic(S, n, seed, penalty=1e5):
    """the operator on a grid of equal spacings in every direction (the
    Jacobian is then symmetric positive definite, as SpdIsai requires)"""
    rng = np.random.default_rng(seed)
    shape = (n, n, n)
    coords = [np.linspace(0.0, 1.0, n)] * 3
    nn = n ** 3
    mu = rng.uniform(1.0, 3.0, nn) * MU0
    flags = (rng.random((n - 1) ** 3) < 0.5).astype(float)
    frac = S.NonlinearMagnetostaticsOperator3D.nodal_fraction(shape, cuda(flags))
    frac_ref = orc.ms_nodal_mu(shape, flags, 0.0)
    op = S.NonlinearMagnetostaticsOperator3D(*coords, cuda(mu), frac, BB, HH, gauge_penalty=penalty)
    return op, orc.Nonlinear(coords, mu, frac_ref, BB, HH, penalty), rng


@pytest.mark.parametrize("shape", [(6, 5, 7), (17, 9, 12)])
@pytest.mark.parametrize("scale", [0.0, 0.8])
def test_assembled_jacobian_matches_the_oracle(S, shape, scale):
    """assemble_matrix_data(model, state) on the device: the oracle's pattern
    exactly (ascending columns, ≤ 23 per row, zero contributions skipped),
    its values within 1e-12, and the product equal to the matrix-free J·v"""
    op, ref, rng = make(S, shape, 31 + sum(shape))
    n = op.size()
    x = scale * rng.standard_normal(n)
    A = op.assemble_jacobian(cuda(x))
    rp, cols, vals = (t.cpu().numpy() for t in A.csr())
    rp_r, cols_r, vals_r = ref.jacobian_csr(x)
    assert np.array_equal(rp, rp_r) and np.array_equal(cols, cols_r)
    assert np.max(np.abs(vals - vals_r)) <= 1e-12 * np.max(np.abs(vals_r))
    v = rng.standard_normal(n)
    jv = A.apply(cuda(v)).cpu().numpy()
    assert rel(jv, ref.jacobian_apply(x, v)) <= 1e-12


def test_newton_with_spd_isai_matches_the_oracle(S):
    """inductor.silpro's setting — Newton with SpdIsai, generated from the
    Jacobian assembled at every step: the oracle's step count (±1), a
    verified solution, fewer inner iterations than Jacobi"""
    op, ref, rng = make_cubic(S, 9, 9)
    b = ref.apply(rng.standard_normal(op.size()))
    tol = 1e-10
    out = {}
    for pc in ("spd_isai", "jacobi"):
        st = S.StrongFormulationSolverSettings(relative_tolerance=tol, max_iterations=20000, preconditioner=pc)
        x, d = S.minimize_strong_formulation_residual(op, cuda(b), settings=st)
        assert d["converged"] and d["verified"], (pc, d)
        r = np.linalg.norm(b - ref.apply(x.cpu().numpy())) / np.linalg.norm(b)
        assert r <= 1.0001 * tol, (pc, r)
        out[pc] = (x.cpu().numpy(), d)
    xr, dr = ref.newton(b, tol, 20000, jacobi="spd_isai")
    assert dr["converged"], dr
    assert abs(out["spd_isai"][1]["outer_iterations"] - dr["newton_steps"]) <= 1, (out["spd_isai"][1], dr)
    assert out["spd_isai"][1]["iterations"] < out["jacobi"][1]["iterations"]
    assert rel(out["spd_isai"][0], xr) <= 1e-5
    print(f"[newton spd_isai] steps {out['spd_isai'][1]['outer_iterations']} (oracle {dr['newton_steps']}), "
          f"inner {out['spd_isai'][1]['iterations']} vs Jacobi {out['jacobi'][1]['iterations']}", flush=True)


def test_nonlinear_solve_rejects_unsupported_settings(S):
    L = __import__("similie_amd._lib", fromlist=["SilError"])
    op, ref, rng = make(S, (6, 5, 7), 2)
    b = cuda(ref.apply(rng.standard_normal(op.size())))
    with pytest.raises(L.SilError, match="SpdIsai preconditioner runs in FP64"):
        S.minimize_strong_formulation_residual(op, b, settings=S.StrongFormulationSolverSettings(
            preconditioner="spd_isai", precision="mixed"))
