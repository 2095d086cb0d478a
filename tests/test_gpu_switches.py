"""The 3D operator's switches on the GPU against the oracle
(magnetostatics_onelab.hpp:796-828, 1893-1941, 2359-2392, 3204-3240,
4930-4935): the linear operator (FP64 and FP32 application, diagonal,
assembled CSR bit for bit), the nonlinear operator and its Jacobian
(application, diagonal, assembly), the CG solve under the divergence gauge,
and the zero right-hand side of the potential criterion."""
import json
from pathlib import Path

import numpy as np
import pytest
import torch

import oracle_lib as orc

pytestmark = pytest.mark.gpu

CURVE = json.loads((Path(__file__).parent / "golden" / "eicore_bh_curve.json").read_text())
MU0 = 4e-7 * np.pi
SWITCHES = [
    dict(divergence_gauge=True),
    dict(divergence_gauge=True, gauge_sign=-1.0),
    dict(y_sign=-1.0),
    dict(y_sign=-1.0, divergence_gauge=True),
]


@pytest.fixture(scope="module")
def S():
    import similie_amd.solvers as S

    return S


def cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def rel(a, b):
    a = a.cpu().numpy() if torch.is_tensor(a) else a
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def grid(shape, rng):
    return [np.concatenate([[0.0], np.cumsum(rng.uniform(0.5, 1.5, n - 1))]) / (n - 1) for n in shape]


def gpu_switches(sw):
    return dict(y_response_sign=sw.get("y_sign", 1.0), divergence_gauge=sw.get("divergence_gauge", False),
                gauge_sign=sw.get("gauge_sign", 1.0))


def host_csr(A):
    rp, cc, vv = A.csr()
    return rp.cpu().numpy().astype(np.int64), cc.cpu().numpy(), vv.cpu().numpy()


@pytest.mark.parametrize("sw", SWITCHES)
@pytest.mark.parametrize("shape", [(6, 5, 7), (40, 33, 21)])
def test_linear_operator_with_switches(S, sw, shape):
    print(f"switches {sw} shape {shape}", flush=True)
    rng = np.random.default_rng(sum(shape) + len(sw))
    coords = grid(shape, rng)
    nn = int(np.prod(shape))
    mu = rng.uniform(1.0, 3.0, nn) * MU0
    op = S.MagnetostaticsOperator3D(*coords, cuda(mu), gauge_penalty=20.0)
    op.set_switches(**gpu_switches(sw))
    x = rng.standard_normal(3 * nn)
    with orc.ms_switches(**sw):
        y_ref = orc.ms_apply(coords, mu, 20.0, x)
        d_ref = orc.ms_diagonal(coords, mu, 20.0)
        erp, ecc, evv = orc.ms_assemble_csr(coords, mu, 20.0)
    y = op.apply(cuda(x))
    torch.cuda.synchronize()
    assert rel(y, y_ref) <= 1e-12
    assert rel(op.apply_f32(cuda(x).float()).double(), y_ref) <= 1e-5
    assert rel(op.diagonal(), d_ref) <= 1e-12
    rp, cc, vv = host_csr(op.assemble())
    np.testing.assert_array_equal(rp, erp)
    np.testing.assert_array_equal(cc, ecc)
    np.testing.assert_array_equal(vv, evv)


@pytest.mark.parametrize("sw", [SWITCHES[0], SWITCHES[2]])
def test_nonlinear_operator_with_switches(S, sw):
    shape = (9, 8, 10)
    print(f"nonlinear switches {sw}", flush=True)
    rng = np.random.default_rng(3)
    coords = grid(shape, rng)
    nn = int(np.prod(shape))
    mu = rng.uniform(1.0, 3.0, nn) * MU0
    ferro = (rng.random(nn) < 0.5).astype(float)
    op = S.NonlinearMagnetostaticsOperator3D(*coords, cuda(mu), cuda(ferro), CURVE["b"], CURVE["h"],
                                             gauge_penalty=1e5)
    op.set_switches(**gpu_switches(sw))
    ref = orc.Nonlinear(coords, mu, ferro, CURVE["b"], CURVE["h"], 1e5)
    x = rng.standard_normal(3 * nn) * 0.02
    v = rng.standard_normal(3 * nn)
    with orc.ms_switches(**sw):
        f_ref = ref.apply(x)
        j_ref = ref.jacobian_apply(x, v)
        d_ref = ref.jacobian_diagonal(x)
        erp, ecc, evv = ref.jacobian_csr(x)
    f = op.apply(cuda(x))
    jv = op.apply_jacobian(cuda(x), cuda(v))
    dg = op.jacobian_diagonal(cuda(x))
    torch.cuda.synchronize()
    assert rel(f, f_ref) <= 1e-12
    assert rel(jv, j_ref) <= 1e-12
    assert rel(dg, d_ref) <= 1e-12
    # entry by entry: where the curve's second-order term vanishes the oracle
    # stores no coupling and the device may hold a round-off one
    import scipy.sparse as sps

    rp, cc, vv = host_csr(op.assemble_jacobian(cuda(x)))
    n = 3 * nn
    got = sps.csr_matrix((vv, cc, rp), shape=(n, n))
    ref_m = sps.csr_matrix((evv, ecc, erp), shape=(n, n))
    assert abs(got - ref_m).max() <= 1e-12 * abs(ref_m).max()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cg_under_the_divergence_gauge(S, precision):
    """the gauge makes the interior operator definite: CG converges to the
    tolerance on the FP64 residual of the oracle's operator"""
    shape = (24, 20, 22)
    print(f"divergence-gauge CG {precision}", flush=True)
    rng = np.random.default_rng(8)
    coords = grid(shape, rng)
    cells = int(np.prod([n - 1 for n in shape]))
    cell_mu = np.where(rng.random(cells) < 0.3, 100.0, 1.0) * MU0
    mu = S.MagnetostaticsOperator3D.nodal_mu(shape, cuda(cell_mu), MU0)
    op = S.MagnetostaticsOperator3D(*coords, mu, gauge_penalty=1.0 / MU0)
    op.set_switches(divergence_gauge=True)
    b = op.rhs(cuda(rng.standard_normal((cells, 3))))
    st = S.StrongFormulationSolverSettings(max_iterations=20000, relative_tolerance=1e-10, precision=precision)
    x, d = S.minimize_strong_formulation_residual(op, b, settings=st)
    assert d["converged"] and d["verified"], d
    with orc.ms_switches(divergence_gauge=True):
        r = b.cpu().numpy() - orc.ms_apply(coords, mu.cpu().numpy(), 1.0 / MU0, x.cpu().numpy())
    assert np.linalg.norm(r) <= 1.0001e-10 * np.linalg.norm(b.cpu().numpy())
    print(f"[divergence gauge cg {precision}] iterations {d['iterations']}", flush=True)


def test_potential_criterion_zeroes_the_rhs(S):
    shape = (7, 6, 8)
    rng = np.random.default_rng(1)
    coords = grid(shape, rng)
    cells = int(np.prod([n - 1 for n in shape]))
    op = S.MagnetostaticsOperator3D(*coords, cuda(np.full(int(np.prod(shape)), MU0)), gauge_penalty=20.0)
    cj = cuda(rng.standard_normal((cells, 3)))
    b = op.rhs(cj)
    assert torch.count_nonzero(b).item() > 0
    op.set_switches(criterion="potential")
    assert torch.count_nonzero(op.rhs(cj)).item() == 0
    op.set_switches(criterion="moments")
    assert torch.equal(op.rhs(cj), b)
