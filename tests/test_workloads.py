"""CPU checks of the synthetic solver workloads (shapes, ranges, symmetry)."""
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
torch = pytest.importorskip("torch")

from similie_amd import workloads as W  # noqa: E402


def test_poisson_bumps():
    x, y, b = W.poisson_bumps()
    assert b.shape == (256, 256) and x[1] - x[0] == pytest.approx(1 / 256)
    assert np.all(b[0] == 0) and np.all(b[:, -1] == 0)
    assert 0.5 <= b.max() <= 2 * 3  # three bumps of amplitude in [0.5, 2]


def test_magnetostatics_inductor_geometry():
    coords, mu, j = W.magnetostatics_inductor(cells=32, core=(8, 8, 16), coil=(10, 12, 8))
    assert len(coords[0]) == 33 and mu.shape == (32 ** 3,) and j.shape == (32 ** 3, 3)
    assert np.count_nonzero(mu > W.MU0 * 2) == 8 * 8 * 16
    assert np.count_nonzero(np.abs(j).sum(1)) == (12 * 12 * 4 - 10 * 10 * 4) * 8
    jj = j.reshape(32, 32, 32, 3)  # [k][j][i][c]
    # the current circulates: antisymmetric under a half turn about the z axis
    assert np.allclose(jj[:, ::-1, ::-1, :2], -jj[:, :, :, :2])
    assert np.all(jj[..., 2] == 0)


def test_gaussian_random_field_statistics():
    gen = torch.Generator().manual_seed(0)
    g = W.gaussian_random_field((128, 96), 1.5, 8.0, gen, device="cpu")
    assert g.shape == (128, 96)
    assert float(g.mean()) == pytest.approx(0.0, abs=1e-12)
    assert float(g.std()) == pytest.approx(1.5, rel=1e-12)
    # correlated: neighbouring values are close compared with the spread
    assert float((g[1:] - g[:-1]).std()) < 0.5 * 1.5


def test_uneven_batch_windows_are_disjoint_and_sized():
    sides, kappa, rhs = W.uneven_poisson_batch(count=40, lo=16, hi=96, device="cpu", field=256)
    assert len(sides) == 40 and sides.min() >= 16 and sides.max() <= 96
    assert kappa.numel() == rhs.numel() == int((sides.astype(np.int64) ** 2).sum())
    assert bool((kappa > 0).all())
    assert float(torch.log(kappa).std()) == pytest.approx(1.0, rel=0.25)
