"""2D magnetostatics on A_z on the GPU against the oracle, through the C ABI:
the linear operator under the three criteria, FP32 application, diagonal,
nodal μ and right-hand side, the CG solve, the nonlinear material (F, its
Jacobian and diagonal, Newton), and the 2D cell post-processing."""
import json
from pathlib import Path

import numpy as np
import pytest
import torch

import oracle_lib as orc

pytestmark = pytest.mark.gpu

MU0 = 4e-7 * np.pi
CURVE = json.loads((Path(__file__).parent / "golden" / "eicore_bh_curve.json").read_text())


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def rel(got, ref):
    g = got.cpu().numpy() if torch.is_tensor(got) else got
    return np.max(np.abs(g - ref)) / max(np.max(np.abs(ref)), 1e-300)


def grid(shape, rng):
    return [np.cumsum(np.concatenate([[rng.uniform(-1, 1)], rng.uniform(0.5, 1.5, n - 1)])) for n in shape]


@pytest.mark.parametrize("shape", [(9, 7), (300, 5), (130, 131)])
@pytest.mark.parametrize("criterion", ["moments", "potential", "potential_and_moments"])
def test_linear_apply_and_diagonal(shape, criterion):
    import similie_amd.solvers as S

    print(f"ms2d linear apply shape={shape} criterion={criterion}", flush=True)
    rng = np.random.default_rng(shape[0] + 7 * shape[1])
    n = shape[0] * shape[1]
    coords = grid(shape, rng)
    mu = rng.uniform(1, 5, n) * MU0
    x = rng.standard_normal(n)
    mu_d, x_d = dev(mu), dev(x)
    op = S.MagnetostaticsOperator2D(*coords, mu_d, criterion=criterion)
    ref = orc.Magnetostatics2D(shape, mu, criterion=criterion)
    y = op.apply(x_d)
    torch.cuda.synchronize()
    assert rel(y, ref.apply(x)) <= 1e-12
    assert rel(op.diagonal(), ref.diagonal()) <= 1e-14
    y32 = op.apply_f32(x_d.float())
    assert rel(y32.double(), ref.apply(x)) <= 1e-5


def test_nodal_mu_and_rhs():
    import similie_amd.solvers as S

    rng = np.random.default_rng(3)
    shape = (37, 29)
    coords = grid(shape, rng)
    cells = 36 * 28
    cmu = rng.uniform(1, 5, cells) * MU0
    jz = rng.standard_normal(cells)
    cmu_d, jz_d = dev(cmu), dev(jz)
    mu = S.MagnetostaticsOperator2D.nodal_mu(shape, cmu_d, MU0)
    torch.cuda.synchronize()
    assert rel(mu, orc.ms2d_nodal_mu(shape, cmu, MU0)) <= 1e-15
    for criterion in ("moments", "potential"):
        op = S.MagnetostaticsOperator2D(*coords, mu, criterion=criterion)
        b = op.rhs(jz_d)
        torch.cuda.synchronize()
        ref = orc.ms2d_rhs(coords, jz, criterion)
        if criterion == "potential":
            assert not torch.any(b).item()
        else:
            assert rel(b, ref) <= 1e-15


def test_linear_solve():
    import similie_amd.solvers as S

    print("ms2d linear CG solve 200x160", flush=True)
    rng = np.random.default_rng(4)
    shape = (200, 160)
    coords = grid(shape, rng)
    cells = 199 * 159
    cmu = np.where(rng.random(cells) < 0.3, 1000 * MU0, MU0)
    jz = np.zeros(cells)
    jz.reshape(159, 199)[60:100, 40:60] = 1e3
    jz.reshape(159, 199)[60:100, 140:160] = -1e3
    cmu_d, jz_d = dev(cmu), dev(jz)
    mu = S.MagnetostaticsOperator2D.nodal_mu(shape, cmu_d, MU0)
    op = S.MagnetostaticsOperator2D(*coords, mu)
    b = op.rhs(jz_d)
    for precision in ("fp64", "mixed"):
        st = S.StrongFormulationSolverSettings(max_iterations=20000, relative_tolerance=1e-10, precision=precision)
        x, d = S.minimize_strong_formulation_residual(op, b, settings=st)
        assert d["converged"], d
        ref = orc.Magnetostatics2D(shape, mu.cpu().numpy())
        r = b.cpu().numpy() - ref.apply(x.cpu().numpy())
        assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(b.cpu().numpy())


def nonlinear_case(shape, rng, level):
    n = shape[0] * shape[1]
    coords = grid(shape, rng)
    ferro = (rng.random(n) < 0.6).astype(float)
    mu = rng.uniform(1, 3, n) * MU0
    x = rng.standard_normal(n) * level
    return coords, ferro, mu, x


@pytest.mark.parametrize("level", [0.05, 0.8])
@pytest.mark.parametrize("criterion", ["moments", "potential_and_moments"])
def test_nonlinear_operator_jacobian_and_diagonal(level, criterion):
    import similie_amd.solvers as S

    print(f"ms2d nonlinear level={level} criterion={criterion}", flush=True)
    rng = np.random.default_rng(int(level * 100))
    shape = (70, 45)
    coords, ferro, mu, x = nonlinear_case(shape, rng, level)
    v = rng.standard_normal(x.size)
    mu_d, fe_d, x_d, v_d = dev(mu), dev(ferro), dev(x), dev(v)
    op = S.NonlinearMagnetostaticsOperator2D(*coords, mu_d, fe_d, CURVE["b"], CURVE["h"], criterion=criterion)
    ref = orc.Magnetostatics2D(shape, mu, ferro, CURVE["b"], CURVE["h"], criterion=criterion)
    fx = op.apply(x_d)
    jv = op.apply_jacobian(x_d, v_d)
    dg = op.jacobian_diagonal(x_d)
    torch.cuda.synchronize()
    assert rel(fx, ref.apply(x)) <= 1e-12
    assert rel(jv, ref.apply(v, x)) <= 1e-12
    assert rel(dg, ref.diagonal(x)) <= 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_newton_matches_the_oracle(precision):
    import similie_amd.solvers as S

    print(f"ms2d Newton {precision}", flush=True)
    shape = (20, 18)
    n = 360
    ferro = np.zeros(n)
    ferro.reshape(18, 20)[6:12, 7:13] = 1.0
    coords = [np.arange(20.0), np.arange(18.0)]
    jz = np.zeros(19 * 17)
    jz.reshape(17, 19)[7:10, 2:5] = 1e5
    jz.reshape(17, 19)[7:10, 14:17] = -1e5
    mu_d, fe_d, jz_d = dev(np.full(n, MU0)), dev(ferro), dev(jz)
    op = S.NonlinearMagnetostaticsOperator2D(*coords, mu_d, fe_d, CURVE["b"], CURVE["h"])
    b = op.rhs(jz_d)
    st = S.StrongFormulationSolverSettings(max_iterations=5000, relative_tolerance=1e-10, precision=precision)
    x, d = S.minimize_strong_formulation_residual(op, b, settings=st)
    ref = orc.Magnetostatics2D(shape, np.full(n, MU0), ferro, CURVE["b"], CURVE["h"])
    xr, dr = ref.newton(b.cpu().numpy(), 1e-10, 5000)
    assert d["converged"] and dr["converged"], (d, dr)
    if precision == "fp64":
        assert d["outer_iterations"] == dr["newton_steps"]
    assert rel(x, xr) <= 1e-7
    r = b.cpu().numpy() - ref.apply(x.cpu().numpy())
    assert np.linalg.norm(r) <= 1.0001e-10 * np.linalg.norm(b.cpu().numpy())


@pytest.mark.parametrize("shape,nonlinear", [((70, 21), True), ((9, 10), False), ((200, 3), True)])
def test_post_process_matches_oracle(shape, nonlinear):
    import similie_amd.solvers as S

    print(f"ms2d post_process shape={shape} nonlinear={nonlinear}", flush=True)
    rng = np.random.default_rng(shape[0])
    coords = grid(shape, rng)
    n = shape[0] * shape[1]
    cells = (shape[0] - 1) * (shape[1] - 1)
    az = rng.standard_normal(n) * 0.7
    cmu = rng.uniform(1, 4, cells) * MU0
    nl = (rng.random(cells) < 0.5).astype(float) if nonlinear else None
    mu_n = dev(np.full(n, MU0))
    if nonlinear:
        op = S.NonlinearMagnetostaticsOperator2D(*coords, mu_n, dev(np.ones(n)), CURVE["b"], CURVE["h"])
    else:
        op = S.MagnetostaticsOperator2D(*coords, mu_n)
    az_d, cmu_d = dev(az), dev(cmu)
    nl_d = dev(nl) if nl is not None else None
    out = op.post_process(az_d, cmu_d, nl_d)
    torch.cuda.synchronize()
    ref = orc.ms2d_post_process(coords, az, cmu, nl, CURVE["b"] if nonlinear else None,
                                CURVE["h"] if nonlinear else None)
    for got, r in zip(out, ref):
        assert rel(got, r) <= 1e-12


def test_mixed_precision_continues_in_fp64_when_it_stalls():
    """On the 2048²-cell μ_r-1000 cross-section the FP32 products cannot
    resolve the residual below ~1e-8; the mixed solve notices the stall and
    finishes in FP64 to the requested FP64-verified tolerance"""
    import similie_amd.solvers as S
    from similie_amd.workloads import magnetostatics_2d_inductor

    print("ms2d mixed-precision stall -> FP64 continuation (2048^2 cells)", flush=True)
    coords, cmu, _, jz = magnetostatics_2d_inductor()
    shape = tuple(len(c) for c in coords)
    mu = S.MagnetostaticsOperator2D.nodal_mu(shape, dev(cmu), MU0)
    op = S.MagnetostaticsOperator2D(*coords, mu)
    b = op.rhs(dev(jz))
    st = S.StrongFormulationSolverSettings(max_iterations=30000, relative_tolerance=1e-10, precision="mixed")
    x, d = S.minimize_strong_formulation_residual(op, b, settings=st)
    assert d["converged"], d
    assert d["true_relative_residual"] <= 1e-10 * (1 + 1e-6)
    r = b - op.apply(x)
    assert (torch.linalg.norm(r) / torch.linalg.norm(b)).item() <= 1.0001e-10


def host_csr(A):
    rp, cc, vv = A.csr()
    return rp.cpu().numpy().astype(np.int64), cc.cpu().numpy(), vv.cpu().numpy()


@pytest.mark.parametrize("shape", [(9, 7), (300, 5), (130, 131)])
@pytest.mark.parametrize("criterion", ["moments", "potential", "potential_and_moments"])
def test_assembled_operator_matches_the_oracle(shape, criterion):
    """assemble_matrix_data of the linear operator: the oracle's CSR arrays
    bit for bit, and its product the matrix-free apply"""
    import similie_amd.solvers as S

    print(f"ms2d assemble shape={shape} criterion={criterion}", flush=True)
    rng = np.random.default_rng(shape[0] * 3 + shape[1])
    n = shape[0] * shape[1]
    coords = grid(shape, rng)
    mu = rng.uniform(1, 5, n) * MU0
    op = S.MagnetostaticsOperator2D(*coords, dev(mu), criterion=criterion)
    A = op.assemble()
    rp, cc, vv = host_csr(A)
    erp, ecc, evv = orc.Magnetostatics2D(shape, mu, criterion=criterion).csr()
    np.testing.assert_array_equal(rp, erp)
    np.testing.assert_array_equal(cc, ecc)
    np.testing.assert_array_equal(vv, evv)
    x = dev(rng.standard_normal(n))
    y_mf, y_asm = op.apply(x), A.apply(x)
    torch.cuda.synchronize()
    assert rel(y_asm, y_mf.cpu().numpy()) <= 1e-13


@pytest.mark.parametrize("level", [0.05, 0.8])
@pytest.mark.parametrize("criterion", ["moments", "potential_and_moments"])
def test_assembled_jacobian_matches_the_oracle(level, criterion):
    """assemble_matrix_data(model, state): the oracle's Jacobian entry by
    entry (relative to the largest, 1e-12), and its product apply_jacobian"""
    import scipy.sparse as sps
    import similie_amd.solvers as S

    print(f"ms2d assemble jacobian level={level} criterion={criterion}", flush=True)
    rng = np.random.default_rng(int(level * 1000) + 5)
    shape = (70, 45)
    coords, ferro, mu, x = nonlinear_case(shape, rng, level)
    n = x.size
    op = S.NonlinearMagnetostaticsOperator2D(*coords, dev(mu), dev(ferro), CURVE["b"], CURVE["h"],
                                             criterion=criterion)
    J = op.assemble_jacobian(dev(x))
    rp, cc, vv = host_csr(J)
    erp, ecc, evv = orc.Magnetostatics2D(shape, mu, ferro, CURVE["b"], CURVE["h"], criterion=criterion).csr(x)
    got = sps.csr_matrix((vv, cc, rp), shape=(n, n))
    ref = sps.csr_matrix((evv, ecc, erp), shape=(n, n))
    assert abs(got - ref).max() <= 1e-12 * abs(ref).max()
    assert np.all(np.diff(rp) <= 7)
    v = dev(rng.standard_normal(n))
    jv, av = op.apply_jacobian(dev(x), v), J.apply(v)
    torch.cuda.synchronize()
    assert rel(av, jv.cpu().numpy()) <= 1e-12


def test_spd_isai_of_the_2d_operator_matches_the_oracle():
    """SpdIsai generated from the linear 2D operator (assembled for the
    purpose) against the oracle's factor of the oracle's matrix; CG with it
    reaches the tolerance in fewer iterations than Jacobi"""
    import similie_amd.solvers as S

    print("ms2d spd_isai 90x70", flush=True)
    rng = np.random.default_rng(21)
    shape = (90, 70)
    n = shape[0] * shape[1]
    coords = grid(shape, rng)
    mu = np.where(rng.random(n) < 0.3, 1000.0, 1.0) * MU0
    op = S.MagnetostaticsOperator2D(*coords, dev(mu))
    grp, gc, gv = host_csr(S.spd_isai(op))
    erp, ecc, evv = orc.spd_isai(*orc.Magnetostatics2D(shape, mu).csr())
    np.testing.assert_array_equal(grp, erp)
    np.testing.assert_array_equal(gc, ecc)
    for r in range(n):
        seg = slice(erp[r], erp[r + 1])
        assert np.max(np.abs(gv[seg] - evv[seg])) <= 1e-12 * np.max(np.abs(evv[seg]))
    b = np.zeros(n)
    b.reshape(shape[1], shape[0])[1:-1, 1:-1] = rng.standard_normal((shape[1] - 2, shape[0] - 2))
    ref = orc.Magnetostatics2D(shape, mu)
    iters = {}
    for pc in ("spd_isai", "jacobi"):
        st = S.StrongFormulationSolverSettings(max_iterations=20000, relative_tolerance=1e-10, preconditioner=pc)
        x, d = S.minimize_strong_formulation_residual(op, dev(b), settings=st)
        assert d["converged"] and d["verified"], (pc, d)
        r = b - ref.apply(x.cpu().numpy())
        assert np.linalg.norm(r) <= 1.0001e-10 * np.linalg.norm(b), pc
        iters[pc] = d["iterations"]
    assert iters["spd_isai"] < iters["jacobi"], iters


def test_newton_with_spd_isai_matches_the_oracle():
    """Newton with SpdIsai of the 2D Jacobian assembled at every step: the
    oracle's step count (±1), a verified solution at the tolerance"""
    import similie_amd.solvers as S

    print("ms2d Newton spd_isai", flush=True)
    shape = (20, 18)
    n = 360
    ferro = np.zeros(n)
    ferro.reshape(18, 20)[6:12, 7:13] = 1.0
    coords = [np.arange(20.0), np.arange(18.0)]
    jz = np.zeros(19 * 17)
    jz.reshape(17, 19)[7:10, 2:5] = 1e5
    jz.reshape(17, 19)[7:10, 14:17] = -1e5
    op = S.NonlinearMagnetostaticsOperator2D(*coords, dev(np.full(n, MU0)), dev(ferro), CURVE["b"], CURVE["h"])
    b = op.rhs(dev(jz))
    st = S.StrongFormulationSolverSettings(max_iterations=5000, relative_tolerance=1e-10, preconditioner="spd_isai")
    x, d = S.minimize_strong_formulation_residual(op, b, settings=st)
    ref = orc.Magnetostatics2D(shape, np.full(n, MU0), ferro, CURVE["b"], CURVE["h"])
    xr, dr = ref.newton(b.cpu().numpy(), 1e-10, 5000, jacobi="spd_isai")
    assert d["converged"] and d["verified"] and dr["converged"], (d, dr)
    assert abs(d["outer_iterations"] - dr["newton_steps"]) <= 1, (d, dr)
    assert rel(x, xr) <= 1e-7
    r = b.cpu().numpy() - ref.apply(x.cpu().numpy())
    assert np.linalg.norm(r) <= 1.0001e-10 * np.linalg.norm(b.cpu().numpy())
