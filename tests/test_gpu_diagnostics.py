"""Air-gap diagnostics on the GPU (magnetostatics_onelab.hpp:4666-4720,
5246-5307, 460-519) against the oracle restatement: random cell fields
over symmetric and asymmetric regions in 3D and 2D (every quantity, the
x-mirror projection included), and a uniform field from the post-processing
of a potential with the closed forms Φ = B·A and traction ½B²/μ0·A."""
import sys
from pathlib import Path

import numpy as np
import pytest
import torch

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import diagnostics_oracle as D  # noqa: E402
from test_oracle_postprocess import uniform_potential  # noqa: E402

pytestmark = pytest.mark.gpu

MU0 = 4e-7 * np.pi


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def compare(got, ref):
    assert got["faces"] == ref["faces"] and got["current_cells"] == ref["current_cells"]
    for key, gk in (("surface_measure", "surface_measure"), ("flux", "flux"),
                    ("traction_magnitude", "traction_magnitude_integral"), ("current", "current")):
        assert abs(got[gk] - ref[key]) <= 1e-12 * max(abs(ref[key]), 1e-300), key
    scale = max(np.max(np.abs(ref["traction"])), 1e-300)
    assert np.max(np.abs(np.array(got["traction"]) - np.array(ref["traction"]))) <= 1e-12 * scale


@pytest.mark.parametrize("symmetric", [True, False])
def test_random_fields_3d(symmetric):
    import similie_amd.solvers as S

    print(f"air-gap diagnostics 3D symmetric={symmetric}", flush=True)
    rng = np.random.default_rng(7 + symmetric)
    # the projection pairs the first and last faces of the (z, x)-sorted
    # list, so it applies only when the faces share one z: one cell layer
    nx, ny, nz = (21, 14, 2) if symmetric else (21, 14, 9)
    x = np.linspace(-1.0, 1.0, nx) if symmetric else np.cumsum(rng.uniform(0.5, 1.5, nx))
    y = np.cumsum(rng.uniform(0.5, 1.5, ny))
    z = np.cumsum(rng.uniform(0.5, 1.5, nz))
    ncx, ncy, ncz = nx - 1, ny - 1, nz - 1
    n = ncx * ncy * ncz
    b, h = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    i = np.arange(n) % ncx
    j = (np.arange(n) // ncx) % ncy
    region = (j >= 4) & (j < 9) & (np.abs(i - (ncx - 1) / 2) < 6)
    region &= (j < 7) | (i < 10) if not symmetric else True
    positive = (i >= 2) & (i < 5) & (j < 6)
    cur = rng.standard_normal((n, 3))
    tags = np.where(region, 7, 1) + np.where(positive & ~region, 2, 0)
    mask = S.diagnostic_cell_mask(tags, diagnostic_region_tags=[7], positive_conductor_tags=[3])
    op = S.MagnetostaticsOperator3D(x, y, z, torch.full((nx * ny * nz,), MU0, dtype=torch.float64, device="cuda"))
    fields = S.CellPostProcessFields(dev(b), dev(h), None, None)
    got = S.air_gap_diagnostics(op, fields, mask, dev(cur), num_turns=4.0)
    ref = D.air_gap_diagnostics((x, y, z), b, h, region, positive & ~region, cur[:, 2])
    compare(got, ref)
    assert ref["faces"] > 0 and ref["current_cells"] > 0
    if symmetric:
        assert got["traction"][0] == 0.0
    assert abs(got["inductance"] - ref["flux"] / (ref["current"] * 4.0)) <= 1e-12 * abs(got["inductance"])


def test_random_fields_2d():
    import similie_amd.solvers as S

    rng = np.random.default_rng(3)
    nx, ny = 33, 20
    x = np.linspace(-2.0, 2.0, nx)
    y = np.cumsum(rng.uniform(0.5, 1.5, ny))
    ncx, ncy = nx - 1, ny - 1
    n = ncx * ncy
    b, h = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    i = np.arange(n) % ncx
    j = np.arange(n) // ncx
    region = (j >= 5) & (j < 8)
    positive = (i < 4) & (j < 4)
    jz = rng.standard_normal(n)
    mask = S.diagnostic_cell_mask(np.where(region, 9, np.where(positive, 3, 1)), [9], [3])
    op = S.MagnetostaticsOperator2D(x, y, torch.full((nx * ny,), MU0, dtype=torch.float64, device="cuda"))
    fields = S.CellPostProcessFields(dev(b), dev(h), None, None)
    got = S.air_gap_diagnostics(op, fields, mask, dev(jz), length_z=0.3, z_value=0.0)
    ref = D.air_gap_diagnostics((x, y), b, h, region, positive, jz, length_z=0.3, z_value=0.0)
    compare(got, ref)
    assert got["traction"][0] == 0.0  # mirror-symmetric faces: projected


def test_uniform_field_closed_form():
    """B = (0, B0, 0) everywhere: Φ = B0·A over the slab's upper face, mean
    traction ½B0²/μ0"""
    import similie_amd.solvers as S

    nx, ny, nz = 17, 12, 9
    coords = [np.linspace(-1.0, 1.0, nx), np.linspace(0.0, 2.0, ny), np.linspace(0.0, 1.0, nz)]
    ncx, ncy, ncz = nx - 1, ny - 1, nz - 1
    n = ncx * ncy * ncz
    mu = torch.full((nx * ny * nz,), MU0, dtype=torch.float64, device="cuda")
    op = S.MagnetostaticsOperator3D(*coords, mu)
    b0 = 0.8
    pot = uniform_potential(coords, [0.0, b0, 0.0]).reshape(-1)
    fields = op.post_process(dev(pot), dev(np.full(n, MU0)))
    j = (np.arange(n) // ncx) % ncy
    mask = S.diagnostic_cell_mask(np.where((j >= 3) & (j < 6), 5, 1), [5], [])
    got = S.air_gap_diagnostics(op, fields, mask, dev(np.zeros((n, 3))))
    area = 2.0 * 1.0
    assert got["faces"] == ncx * ncz
    assert abs(got["flux"] - b0 * area) <= 1e-9 * b0 * area
    p = 0.5 * b0 * b0 / MU0
    assert abs(got["mean_traction_magnitude"] - p) <= 1e-9 * p
    np.testing.assert_allclose(got["traction"], [0.0, p * area, 0.0], rtol=1e-9, atol=1e-6 * p)
