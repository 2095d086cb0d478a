"""The air-gap diagnostics oracle (magnetostatics_onelab.hpp:4666-4720,
5246-5307, 460-519) against closed forms: a uniform field over a slab gives
Φ = B·A and the Maxwell traction ½B²/μ0 on its upper face, the x-mirror
projection applies only to mirror-symmetric faces, and the conductor
current is J times its cross-section."""
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import diagnostics_oracle as D  # noqa: E402

MU0 = 4e-7 * np.pi


def uniform(ncell, b0):
    b = np.zeros((ncell, 3))
    b[:, 1] = b0
    return b, b / MU0


def test_uniform_field_over_a_slab_3d():
    x = np.linspace(-2.0, 2.0, 9)
    y = np.linspace(0.0, 3.0, 7)
    z = np.linspace(0.0, 1.5, 4)
    ncx, ncy, ncz = 8, 6, 3
    n = ncx * ncy * ncz
    b, h = uniform(n, 0.7)
    j = (np.arange(n) // ncx) % ncy
    region = (j >= 2) & (j < 4)
    r = D.air_gap_diagnostics((x, y, z), b, h, region, np.zeros(n, bool), np.zeros(n))
    area = 4.0 * 1.5
    assert r["faces"] == ncx * ncz
    assert abs(r["surface_measure"] - area) <= 1e-12
    assert abs(r["flux"] - 0.7 * area) <= 1e-12
    p = 0.5 * 0.7 ** 2 / MU0
    np.testing.assert_allclose(r["traction"], [0.0, p * area, 0.0], rtol=1e-12, atol=1e-9)
    assert abs(r["traction_magnitude"] - p * area) <= 1e-9 * p * area


def test_mirror_projection_needs_symmetric_faces():
    x = np.array([-1.5, -0.5, 0.5, 1.5])
    y = np.array([0.0, 1.0, 2.0])
    ncx, ncy = 3, 2
    n = ncx * ncy
    rng = np.random.default_rng(2)
    b = rng.standard_normal((n, 3))
    h = rng.standard_normal((n, 3))
    region = np.zeros(n, bool)
    region[:ncx] = True  # the lower row: faces at x = −1, 0, 1
    r = D.air_gap_diagnostics((x, y), b, h, region, np.zeros(n, bool), np.zeros(n), length_z=2.0)
    t = [D.traction_on_positive_y_face(b[i], h[i]) for i in range(3)]
    assert r["traction"][0] == 0.0
    np.testing.assert_allclose(r["traction"][1], 2.0 * (t[0][1] + t[2][1]) + 2.0 * t[1][1], rtol=1e-14)
    # drop the middle face: still symmetric
    region[1] = False
    r = D.air_gap_diagnostics((x, y), b, h, region, np.zeros(n, bool), np.zeros(n), length_z=2.0)
    assert r["traction"][0] == 0.0 and r["faces"] == 2
    # an unpaired face: the raw integral stays
    region[0] = False
    r = D.air_gap_diagnostics((x, y), b, h, region, np.zeros(n, bool), np.zeros(n), length_z=2.0)
    np.testing.assert_allclose(r["traction"], [2.0 * t[2][0], 2.0 * t[2][1], 2.0 * t[2][2]], rtol=1e-14)


def test_current_is_j_times_the_cross_section():
    x = np.linspace(0.0, 4.0, 5)
    y = np.linspace(0.0, 3.0, 4)
    z = np.linspace(0.0, 4.0, 5)
    ncx, ncy, ncz = 4, 3, 4
    n = ncx * ncy * ncz
    i = np.arange(n) % ncx
    k = np.arange(n) // (ncx * ncy)
    pos = (i == 1)
    jz = np.where(pos, 5.0, 0.0)
    r = D.air_gap_diagnostics((x, y, z), np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n, bool), pos, jz)
    assert r["current_cells"] == ncy  # the layer k = ncz // 2 only
    assert abs(r["current"] - 5.0 * 1.0 * 3.0) <= 1e-12
    assert np.all(k[pos & (k == ncz // 2)] == 2)
