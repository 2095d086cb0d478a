"""Cell post-processing on the GPU (B, H, Maxwell stress, xy-slice force
density) against the oracle, through the C ABI."""
import json
from pathlib import Path

import numpy as np
import pytest
import torch

import oracle_lib as orc
from test_oracle_postprocess import axes, ncells, uniform_potential

pytestmark = pytest.mark.gpu

MU0 = 4e-7 * np.pi
CURVE = json.loads((Path(__file__).parent / "golden" / "eicore_bh_curve.json").read_text())
NAMES = ("magnetic_induction", "magnetic_field", "maxwell_stress", "force_density")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def operator(coords, nonlinear):
    import similie_amd.solvers as S

    shape = tuple(len(c) for c in coords)
    mu_n = torch.full((int(np.prod(shape)),), MU0, dtype=torch.float64, device="cuda")
    if nonlinear:
        frac = torch.ones_like(mu_n)
        return S.NonlinearMagnetostaticsOperator3D(*coords, mu_n, frac, CURVE["b"], CURVE["h"])
    return S.MagnetostaticsOperator3D(*coords, mu_n)


def check(got, ref, what):
    g = got.cpu().numpy()
    scale = max(np.max(np.abs(ref)), 1e-300)
    err = np.max(np.abs(g - ref)) / scale if ref.size else 0.0
    assert err <= 1e-12, f"{what}: relative error {err:.3e}"


@pytest.mark.parametrize("shape,nonlinear", [((70, 21, 5), True), ((66, 18, 4), False), ((9, 10, 3), True),
                                             ((2, 7, 3), False), ((130, 4, 3), True)])
def test_post_process_matches_oracle(shape, nonlinear):
    print(f"post_process parity shape={shape} nonlinear={nonlinear}", flush=True)
    rng = np.random.default_rng(sum(shape))
    coords = axes(shape, rng)
    n = ncells(coords)
    pot = rng.standard_normal(3 * int(np.prod(shape))) * 0.7
    mu = rng.uniform(1.0, 4.0, n) * MU0
    nl = (rng.random(n) < 0.5).astype(float) if nonlinear else None
    op = operator(coords, nonlinear)
    out = op.post_process(dev(pot), dev(mu), dev(nl) if nl is not None else None)
    torch.cuda.synchronize()
    ref = orc.ms_post_process(coords, pot, mu, nl, CURVE["b"] if nonlinear else None,
                              CURVE["h"] if nonlinear else None)
    for name, got, r in zip(NAMES, out, ref):
        check(got, r, name)


def test_post_process_outputs_are_optional():
    rng = np.random.default_rng(5)
    coords = axes((40, 12, 4), rng)
    n = ncells(coords)
    pot = rng.standard_normal(3 * 40 * 12 * 4)
    mu = rng.uniform(1.0, 4.0, n) * MU0
    op = operator(coords, False)
    import similie_amd._lib as L

    force = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    pot_d, mu_d = dev(pot), dev(mu)  # held: a temporary's memory could be reused before the launch
    L.check(L.lib().sil_magnetostatics3d_post_process(op.handle, L.ptr(pot_d), L.ptr(mu_d), None, None, None,
                                                      None, L.ptr(force), None))
    torch.cuda.synchronize()
    check(force, orc.ms_post_process(coords, pot, mu)[3], "force_density alone")


def test_nonlinear_cells_need_a_nonlinear_operator():
    import similie_amd._lib as L

    coords = [np.arange(5, dtype=float)] * 3
    op = operator(coords, False)
    z = torch.zeros(3 * 125, dtype=torch.float64, device="cuda")
    with pytest.raises(L.SilError, match="nonlinear operator"):
        op.post_process(z, torch.ones(64, dtype=torch.float64, device="cuda"),
                        torch.ones(64, dtype=torch.float64, device="cuda"))


def test_uniform_field_at_full_size():
    """192³ cells: a uniform B0 is recovered in every cell and, on the
    uniform grid, the force density vanishes off the first row and column
    of every slice"""
    print("post_process uniform field 192^3 cells", flush=True)
    shape = (193, 193, 193)
    coords = [np.arange(n, dtype=float) * 0.01 for n in shape]
    B0 = np.array([0.3, -0.7, 1.1])
    pot = uniform_potential(coords, B0)
    n = ncells(coords)
    op = operator(coords, True)
    nl = torch.zeros(n, dtype=torch.float64, device="cuda")
    nl[: n // 2] = 1.0
    mu = torch.full((n,), 3 * MU0, dtype=torch.float64, device="cuda")
    out = op.post_process(dev(pot), mu, nl)
    torch.cuda.synchronize()
    B = out.magnetic_induction
    assert torch.max(torch.abs(B - torch.from_numpy(B0).cuda())).item() <= 1e-12 * 1.1
    F = out.force_density.view(192, 192, 192, 3)
    scale = torch.max(torch.abs(out.maxwell_stress)).item() / 0.01 ** 2
    lin = F[96:, 1:, 1:, :]  # linear half: uniform stress
    assert torch.max(torch.abs(lin)).item() <= 1e-11 * scale
    assert torch.min(torch.abs(F[96:, 0, 1:, 1])).item() > 1e-3 * scale


def test_solve_then_post_process():
    """the tagged inductor solved on the GPU, post-processed on the GPU and
    by the oracle from the same potential"""
    import similie_amd.io as IO
    import similie_amd.solvers as S
    from test_gpu_io_pipeline import tagged_inductor

    print("post_process after a nonlinear solve (24^3 cells)", flush=True)
    n = 24
    coords = [np.arange(n + 1, dtype=float)] * 3
    tags = tagged_inductor(n)
    mu, cur, flags = IO.magnetostatics_cell_inputs(
        tags, MU0, core_mu=1000 * MU0, current_density_x=2e3, current_density_z=2e3, magnetic_material_tags=[2],
        positive_x_tags=[5], negative_x_tags=[6], positive_z_tags=[3], negative_z_tags=[4], nonlinear=True)
    shape = (n + 1,) * 3
    mu_n = S.MagnetostaticsOperator3D.nodal_mu(shape, dev(mu), MU0)
    frac = S.NonlinearMagnetostaticsOperator3D.nodal_fraction(shape, dev(flags))
    op = S.NonlinearMagnetostaticsOperator3D(*coords, mu_n, frac, CURVE["b"], CURVE["h"])
    x, d = S.minimize_strong_formulation_residual(op, op.rhs(dev(cur)),
                                                  settings=S.StrongFormulationSolverSettings(max_iterations=20000, relative_tolerance=1e-8))
    assert d["converged"], d
    out = op.post_process(x, dev(mu), dev(flags))
    torch.cuda.synchronize()
    ref = orc.ms_post_process(coords, x.cpu().numpy(), mu, flags, CURVE["b"], CURVE["h"])
    for name, got, r in zip(NAMES, out, ref):
        check(got, r, name)
    # the core carries the flux: |B| there well above the air's
    B = out.magnetic_induction.cpu().numpy()
    assert np.median(np.linalg.norm(B[tags == 2], axis=1)) > 10 * np.median(np.linalg.norm(B[tags == 1], axis=1))
