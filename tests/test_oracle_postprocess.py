"""Pins of the oracle's cell post-processing (CPU): the cell induction of a
uniform field, the constitutive laws, the Maxwell stress and the xy-slice
force density, each against what the mathematics fixes.  The reference has
no test of its post-processing, so these are the pins."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as orc

MU0 = 4e-7 * np.pi
CURVE = json.loads((Path(__file__).parent / "golden" / "eicore_bh_curve.json").read_text())


def axes(shape, rng, uniform=False):
    if uniform:
        return [np.arange(n, dtype=float) * 0.25 for n in shape]
    return [np.cumsum(np.concatenate([[rng.uniform(-1, 1)], rng.uniform(0.5, 1.5, n - 1)])) for n in shape]


def edge_cochain(coords, field):
    """the potential as a one-form: ∫ A·dl along each forward edge (exact
    at the edge midpoint for a linear field), [node][3], x fastest"""
    nx, ny, nz = (len(c) for c in coords)
    out = np.zeros((nz, ny, nx, 3))
    for c in range(3):
        w = np.diff(coords[c])
        mid = [coords[0], coords[1], coords[2]]
        mid[c] = 0.5 * (coords[c][1:] + coords[c][:-1])
        Z, Y, X = np.meshgrid(mid[2], mid[1], mid[0], indexing="ij")
        A = field(X, Y, Z)[c]
        shp = [1, 1, 1]
        shp[2 - c] = len(w)
        val = A * w.reshape(shp)
        sl = [slice(None)] * 3
        sl[2 - c] = slice(0, len(w))
        out[tuple(sl) + (c,)] = val
    return out.reshape(-1)


def uniform_potential(coords, B0):
    B0 = np.asarray(B0, dtype=float)

    def field(X, Y, Z):  # A = ½ B0 × r
        return (0.5 * (B0[1] * Z - B0[2] * Y), 0.5 * (B0[2] * X - B0[0] * Z), 0.5 * (B0[0] * Y - B0[1] * X))

    return edge_cochain(coords, field)


def ncells(coords):
    return int(np.prod([len(c) - 1 for c in coords]))


def centre_spacing(c):
    m = 0.5 * (c[1:] + c[:-1])
    if len(m) == 1:
        return np.zeros(1)
    h = np.empty(len(m))
    h[:-1] = np.diff(m)
    h[-1] = m[-1] - m[-2]
    return h


def force_closed_form(coords, stress):
    """−δ of the three stress one-forms on each xy slice of cell centres:
    f = [(hy/hx)ω_x − (hy/hx)ω_x|_(i−1) + (hx/hy)ω_y − (hx/hy)ω_y|_(j−1)] / (hx hy),
    with zeros outside the slice"""
    mx, my, mz = (len(c) - 1 for c in coords)
    hx, hy = centre_spacing(coords[0]), centre_spacing(coords[1])
    T = stress.reshape(mz, my, mx, 6)
    HX, HY = hx[None, None, :], hy[None, :, None]
    out = np.zeros((mz, my, mx, 3))
    with np.errstate(divide="ignore", invalid="ignore"):
        for f, (a, b) in enumerate(((0, 3), (3, 1), (4, 5))):
            ex = (HY / HX) * T[..., a]
            ey = (HX / HY) * T[..., b]
            g = ex.copy()
            g[:, :, 1:] -= ex[:, :, :-1]
            g += ey
            g[:, 1:, :] -= ey[:, :-1, :]
            out[..., f] = g / (HX * HY)
    return out.reshape(-1, 3)


@pytest.mark.parametrize("B0", [(0.3, -0.7, 1.1), (0.0, 0.0, 2.0), (-1.5, 0.25, 0.0)])
def test_uniform_field_gives_uniform_induction(B0):
    rng = np.random.default_rng(1)
    coords = axes((7, 6, 5), rng)
    pot = uniform_potential(coords, B0)
    B, H, T, F = orc.ms_post_process(coords, pot, np.full(ncells(coords), MU0))
    np.testing.assert_allclose(B, np.broadcast_to(B0, B.shape), rtol=0, atol=1e-12 * np.max(np.abs(B0)))


def test_single_edge_moments_signs():
    """a unit A_x on the edge (i, j, k)→(i+1, j, k) enters B_y of the cells
    (i, j, k−1) and (i, j, k) and B_z of the cells (i, j−1, k) and (i, j, k)
    with the orientation of the face circulations"""
    coords = [np.arange(4, dtype=float)] * 3
    pot = np.zeros(3 * 64)
    i, j, k = 1, 1, 1
    pot[3 * (i + 4 * (j + 4 * k)) + 0] = 1.0
    B, _, _, _ = orc.ms_post_process(coords, pot, np.full(27, MU0))
    B = B.reshape(3, 3, 3, 3)  # [k][j][i][m]
    assert B[k, j, i, 1] == -1.0 and B[k - 1, j, i, 1] == 1.0
    assert B[k, j, i, 2] == 1.0 and B[k, j - 1, i, 2] == -1.0
    assert np.count_nonzero(B) == 4


def test_linear_and_nonlinear_constitutive_laws():
    rng = np.random.default_rng(2)
    coords = axes((6, 5, 5), rng)
    n = ncells(coords)
    k = 7
    b_k = CURVE["b"][k]
    direction = np.array([0.48, -0.6, 0.64])
    pot = uniform_potential(coords, b_k * direction)
    mu = rng.uniform(1.0, 5.0, n) * MU0
    nl = (rng.random(n) < 0.5).astype(float)
    B, H, T, _ = orc.ms_post_process(coords, pot, mu, nl, CURVE["b"], CURVE["h"])
    lin = nl == 0
    np.testing.assert_allclose(H[lin], B[lin] / mu[lin, None], rtol=1e-15)
    hn = np.linalg.norm(H[~lin], axis=1)
    np.testing.assert_allclose(hn, CURVE["h"][k], rtol=1e-12)
    np.testing.assert_allclose(H[~lin] / hn[:, None], direction[None, :].repeat((~lin).sum(), 0), atol=1e-12)
    # zero induction: zero field on a nonlinear cell
    _, H0, T0, F0 = orc.ms_post_process(coords, np.zeros_like(pot), mu, np.ones(n), CURVE["b"], CURVE["h"])
    assert not np.any(H0) and not np.any(T0) and not np.any(F0)


def test_maxwell_stress_closed_form():
    rng = np.random.default_rng(3)
    coords = axes((5, 6, 4), rng)
    n = ncells(coords)
    b = 1.7
    pot = uniform_potential(coords, (0.0, 0.0, b))
    B, H, T, _ = orc.ms_post_process(coords, pot, np.full(n, MU0))
    p = b * b / (2 * MU0)
    expect = np.array([-p, -p, p, 0.0, 0.0, 0.0])
    np.testing.assert_allclose(T, np.broadcast_to(expect, T.shape), rtol=0, atol=1e-12 * p)
    # trace = −½ B·H on a general field
    pot = rng.standard_normal(3 * int(np.prod([len(c) for c in coords])))
    B, H, T, _ = orc.ms_post_process(coords, pot, rng.uniform(1, 3, n) * MU0)
    bh = np.sum(B * H, axis=1)
    np.testing.assert_allclose(T[:, 0] + T[:, 1] + T[:, 2], -0.5 * bh, rtol=1e-12, atol=1e-12 * np.max(np.abs(bh)))
    np.testing.assert_allclose(T[:, 3:], np.stack([B[:, 0] * H[:, 1], B[:, 0] * H[:, 2], B[:, 1] * H[:, 2]], 1))


@pytest.mark.parametrize("shape,uniform", [((9, 8, 4), False), ((6, 7, 3), True), ((2, 6, 3), False)])
def test_force_density_is_the_slice_codifferential(shape, uniform):
    """the oracle's force (through the general DEC codifferential with a
    stored identity metric) equals the closed form of −δ on the centre grid"""
    rng = np.random.default_rng(4)
    coords = axes(shape, rng, uniform)
    n = ncells(coords)
    pot = rng.standard_normal(3 * int(np.prod(shape)))
    nl = (rng.random(n) < 0.4).astype(float)
    _, _, T, F = orc.ms_post_process(coords, pot, rng.uniform(1, 3, n) * MU0, nl, CURVE["b"], CURVE["h"])
    ref = force_closed_form(coords, T)
    if shape[0] == 2:  # one cell along x: every ⋆ coefficient is degenerate
        ref[:] = 0.0
    np.testing.assert_allclose(F, ref, rtol=0, atol=1e-12 * max(np.max(np.abs(ref)), 1e-300))


def test_uniform_stress_has_no_interior_force():
    coords = axes((8, 7, 4), None, uniform=True)
    n = ncells(coords)
    pot = uniform_potential(coords, (0.4, 0.9, -0.3))
    _, _, T, F = orc.ms_post_process(coords, pot, np.full(n, MU0))
    F = F.reshape(3, 6, 7, 3)
    scale = np.max(np.abs(T)) / 0.25 ** 2
    assert np.max(np.abs(F[:, 1:, 1:, :])) <= 1e-12 * scale
    # the first row and column see the zero outside the slice
    assert np.min(np.abs(F[:, 0, 1:, 1])) > 1e-3 * scale
