"""Pins the CPU oracle before anything is compared against it.

Every check here is either a known answer from the reference's own test
suite (tests/exterior/*.cpp) or a property the mathematics fixes (closed
forms, transposes, Poincaré, cancellations), so that a dropped term, a
wrong sign or a transposed index in the oracle fails one of them.
"""
import json
import math
from itertools import combinations
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as orc

GOLDEN = Path(__file__).with_name("golden")


def harness_input(ndim, npts, degree):
    """tests/exterior/derivative.cpp:15-99: component i holds i+1 everywhere
    and i+2 at node (1,...,1)."""
    nc = orc.binom(ndim, degree)
    arr = np.zeros((npts,) * ndim + (nc,))
    for i in range(nc):
        arr[..., i] = i + 1.0
        arr[(1,) * ndim + (i,)] = i + 2.0
    return arr


@pytest.mark.parametrize("case", json.loads((GOLDEN / "derivative_known_answers.json").read_text())["cases"],
                         ids=lambda c: f"{c['test']}-{'first' if c['index_first'] else 'last'}")
def test_coboundary_reference_known_answers(case):
    n, npts = case["ndim"], case["points_per_axis"]
    k = case["in_degree"]
    arr = harness_input(n, npts, k)
    out_ext = [npts if case["out_degree"] == 1 else npts - 1] * n
    out = orc.coboundary(n, k, [npts] * n, arr, out_ext)
    assert np.all(np.isfinite(out))
    for e in case["expects"]:
        comp = orc.combo_index(n, k + 1, e["axes_mask"])
        assert out[tuple(e["node"]) + (comp,)] == e["value"], e


def test_coboundary_of_boundary_cochain():
    """tests/exterior/exterior.cpp:543-564: the XY 2-simplex of a 4D
    (T,X,Y,Z) grid, boundary values 5, 8, 3, 2 in boundary() order → 4."""
    faces = orc.boundary(4, 0b0110)
    vals = [5.0, 8.0, 3.0, 2.0]
    assert sum((-1.0 if neg else 1.0) * v for (_, _, neg), v in zip(faces, vals)) == 4.0


@pytest.mark.parametrize("n", [2, 3, 4])
def test_poincare_boundary_of_boundary_vanishes(n):
    """tests/exterior/exterior.cpp:433-496: ∂∂ = 0 for every simplex."""
    for k in range(2, n + 1):
        for axes in combinations(range(n), k):
            mask = sum(1 << a for a in axes)
            acc = {}
            for off, vm, neg in orc.boundary(n, mask):
                for off2, vm2, neg2 in orc.boundary(n, vm):
                    key = (tuple(o + p for o, p in zip(off, off2)), vm2)
                    acc[key] = acc.get(key, 0) + (-1 if neg ^ neg2 else 1)
            assert all(v == 0 for v in acc.values())


def closed_form_d(arr, n, k, out_ext):
    """(dω)_A(n) = Σ_m (-1)^m [ω_{A∖a_m}(clamp(n+e_{a_m})) − ω_{A∖a_m}(n)]."""
    ext = arr.shape[:n]
    lo = [sum(1 << a for a in c) for c in combinations(range(n), k)]
    up = [sum(1 << a for a in c) for c in combinations(range(n), k + 1)]
    out = np.zeros(tuple(out_ext) + (len(up),))
    for ci, A in enumerate(up):
        axes = [a for a in range(n) if A >> a & 1]
        for m, a in enumerate(axes):
            li = lo.index(A & ~(1 << a))
            idx_n = tuple(slice(0, e) for e in out_ext)
            shifted = []
            for ax in range(n):
                r = np.arange(out_ext[ax]) + (1 if ax == a else 0)
                shifted.append(np.minimum(r, ext[ax] - 1))
            fwd = arr[np.ix_(*shifted)][..., li]
            out[..., ci] += (-1) ** m * (fwd - arr[idx_n + (li,)])
    return out


@pytest.mark.parametrize("n,k", [(1, 0), (2, 0), (2, 1), (3, 0), (3, 1), (3, 2), (4, 1), (4, 2), (4, 3)])
def test_coboundary_matches_closed_form_and_dd_zero(n, k):
    rng = np.random.default_rng(10 * n + k)
    ext = [5, 4, 6, 3][:n]
    arr = rng.integers(-1000, 1001, size=tuple(ext) + (orc.binom(n, k),)).astype(np.float64)
    out = orc.coboundary(n, k, ext, arr, ext)
    np.testing.assert_array_equal(out, closed_form_d(arr, n, k, ext))
    if k + 1 < n:
        dd = orc.coboundary(n, k + 1, ext, out, ext)
        inner = tuple(slice(0, e - 1) for e in ext)
        assert np.all(dd[inner] == 0.0)


@pytest.mark.parametrize("n,k", [(1, 0), (2, 0), (2, 1), (3, 0), (3, 1), (3, 2)])
def test_transposed_coboundary_matches_backward_closed_form(n, k):
    """(dᵀη)_A(n) = Σ_m (−1)^m [η_{A∖a_m}(n) − η_{A∖a_m}(n−e_{a_m})], η ≡ 0 outside."""
    rng = np.random.default_rng(7 + n + k)
    ext = [6, 5, 4][:n]
    x = rng.standard_normal(tuple(ext) + (orc.binom(n, k),))
    tx = orc.transposed_coboundary(n, k, ext, x, ext)
    lo = [sum(1 << a for a in c) for c in combinations(range(n), k)]
    up = [sum(1 << a for a in c) for c in combinations(range(n), k + 1)]
    ref = np.zeros_like(tx)
    xp = np.pad(x, [(1, 0)] * n + [(0, 0)])
    for ci, A in enumerate(up):
        axes = [a for a in range(n) if A >> a & 1]
        for m, a in enumerate(axes):
            li = lo.index(A & ~(1 << a))
            cur = xp[tuple(slice(1, None) for _ in range(n)) + (li,)]
            back = xp[tuple(slice(0, -1) if ax == a else slice(1, None) for ax in range(n)) + (li,)]
            ref[..., ci] += (-1) ** m * (cur - back)
    np.testing.assert_allclose(tx, ref, rtol=0, atol=1e-13)


@pytest.mark.parametrize("n,k", [(1, 1), (2, 1), (2, 2), (3, 1), (3, 2), (3, 3)])
def test_codifferential_is_adjoint_of_d(n, k):
    """(dα, β)_k = (α, δβ)_{k−1} with (α, β)_k = Σ_nodes Σ_I α_I β_I V_J/V_I,
    for α, β supported away from the boundary.  This ties d, dᵀ, both stars
    and the sign rule of δ together."""
    rng = np.random.default_rng(40 + 10 * n + k)
    ext = [7, 6, 6][:n]
    axes = [np.cumsum(np.concatenate([[0.0], rng.uniform(0.5, 1.5, e - 1)])) for e in ext]
    g = rng.uniform(0.5, 2.0, size=(int(np.prod(ext)), n))
    geom = orc.Geometry(ext, axes=axes, metric_kind=orc.METRIC_DIAGONAL, metric=g)
    inner = tuple(slice(2, e - 2) for e in ext)
    alpha = np.zeros(tuple(ext) + (orc.binom(n, k - 1),))
    beta = np.zeros(tuple(ext) + (orc.binom(n, k),))
    alpha[inner] = rng.standard_normal(alpha[inner].shape)
    beta[inner] = rng.standard_normal(beta[inner].shape)

    def weights(deg):
        w = np.zeros(tuple(ext) + (orc.binom(n, deg),))
        for lin in range(int(np.prod(ext))):
            node = list(np.unravel_index(lin, ext))
            for ci, src in enumerate(combinations(range(n), deg)):
                sm = sum(1 << a for a in src)
                w[tuple(node) + (ci,)] = abs(orc.hodge_coefficient(geom, node, sm, ((1 << n) - 1) & ~sm))
        return w

    da = orc.coboundary(n, k - 1, ext, alpha, ext)
    db = orc.codifferential(geom, k, beta, ext)
    lhs = np.sum(da * beta * weights(k))
    rhs = np.sum(alpha * db * weights(k - 1))
    assert math.isclose(lhs, rhs, rel_tol=1e-12, abs_tol=1e-12)


def spd_metric_3d():
    return np.array([4.0, 1.0, 2.0, 5.0, 3.0, 6.0])  # XX XY XZ YY YZ ZZ


def test_discrete_hodge_star_reference_golden():
    """tests/exterior/hodge_star.cpp:65-181: ⋆(3 dx∧dy) = 1.5·√(6/19) dz and
    the dual star brings it back to 3, with metric [[4,1,2],[1,5,3],[2,3,6]]
    on a 2×2×2 grid of spacing 2."""
    ext = [2, 2, 2]
    metric = np.tile(spd_metric_3d(), (8, 1))
    geom = orc.Geometry(ext, axes=[np.array([0.0, 2.0])] * 3, metric_kind=orc.METRIC_SYMMETRIC, metric=metric)
    for chain in (False, True):
        h = orc.fill_hodge_star(geom, 2, use_chain=chain)  # [node][9][3]
        form = np.zeros(9)
        form[0 * 3 + 1] = 3.0  # natural (X,Y)
        form[1 * 3 + 0] = -3.0  # natural (Y,X)
        dual = np.einsum("...ft,f->...t", h, form)
        assert np.allclose(dual[..., 2], 1.5 * math.sqrt(6.0 / 19.0), rtol=4e-16 * 4, atol=0)
        assert np.allclose(dual[..., :2], 0.0, atol=1e-15)
        h2 = orc.fill_hodge_star(geom, 1, use_chain=chain)  # [node][3][3]
        back = np.einsum("...ft,...f->...t", h2, dual)
        assert np.allclose(back[..., 0], 3.0, rtol=1e-15)
        assert np.allclose(back[..., 1:], 0.0, atol=1e-15)


def test_continuous_hodge_star_reference_golden():
    """tests/exterior/hodge_star.cpp:183-265: the continuous ⋆ applied to
    3 dx∧dy gives 18·√70/70 (two natural orderings of the source pair)."""
    metric = spd_metric_3d()
    v = orc.continuous_hodge_value([1, 1, 1], orc.METRIC_SYMMETRIC, metric, [0, 0, 0], 0b011, 0b100)
    assert math.isclose(2 * 3.0 * v, 18.0 * math.sqrt(70.0) / 70.0, rel_tol=4e-16 * 4)


def skew_positions(ext, xi_eta):
    pos = np.zeros(tuple(ext) + (2,))
    for i in range(ext[0]):
        for j in range(ext[1]):
            pos[i, j] = xi_eta(2.0 * i, 2.0 * j)
    return pos


@pytest.mark.parametrize("k,form,expected", [(0, [3.5], [3.5]), (1, [3.0, -1.0], [6.0, 2.0]), (2, [4.0], [32.0])])
def test_reduction_reference_golden(k, form, expected):
    """tests/exterior/reduction_and_reconstruction.cpp:231-285: positions
    x = ξ+η, y = 2η on a 2×2 grid of spacing 2."""
    pos = skew_positions([2, 2], lambda xi, eta: (xi + eta, 2.0 * eta))
    geom = orc.Geometry([2, 2], nodal=pos)
    arr = np.tile(np.array(form), (2, 2, 1))
    red = orc.reduction(geom, k, arr)
    assert red[0, 0].tolist() == expected


@pytest.mark.parametrize("k,form", [(0, [3.5]), (1, [3.0, -1.0]), (2, [4.0])])
def test_reconstruction_round_trip_reference_golden(k, form):
    """tests/exterior/reduction_and_reconstruction.cpp:287-329: positions
    (2ξ, 3η); reconstruction after reduction is the identity."""
    pos = skew_positions([2, 2], lambda xi, eta: (2.0 * xi, 3.0 * eta))
    geom = orc.Geometry([2, 2], nodal=pos)
    arr = np.tile(np.array(form), (2, 2, 1))
    back = orc.reconstruction(geom, k, orc.reduction(geom, k, arr))
    np.testing.assert_allclose(back[0, 0], form, rtol=1e-15)


def random_spd(rng, n):
    a = rng.standard_normal((n, n))
    m = a @ a.T + n * np.eye(n)
    return m


@pytest.mark.parametrize("n", [2, 3])
def test_hodge_chain_equals_volume_ratio_on_general_geometry(n):
    """The reference computes ⋆ through reconstruction → continuous star →
    circumcentric reduction; the closed form sign(I‖J)·V_J/V_I must agree
    with that chain for arbitrary SPD metrics and skewed positions."""
    rng = np.random.default_rng(3 + n)
    ext = [3] * n
    nn = 3 ** n
    sym = []
    for _ in range(nn):
        m = random_spd(rng, n)
        sym.append([m[i, j] for i in range(n) for j in range(i, n)])
    base = np.stack(np.meshgrid(*[np.arange(3.0)] * n, indexing="ij"), axis=-1)
    pos = base + 0.15 * rng.standard_normal(base.shape)
    geom = orc.Geometry(ext, nodal=pos.reshape(-1, n), metric_kind=orc.METRIC_SYMMETRIC, metric=np.array(sym))
    for node in [[0] * n, [1] * n, [2] * n, [0, 2, 1][:n]]:
        for k in range(n + 1):
            for src in combinations(range(n), k):
                sm = sum(1 << a for a in src)
                tm = ((1 << n) - 1) & ~sm
                a = orc.hodge_coefficient(geom, node, sm, tm)
                b = orc.hodge_coefficient(geom, node, sm, tm, chain=True)
                assert math.isclose(a, b, rel_tol=1e-12), (node, src, a, b)


@pytest.mark.parametrize("n", [2, 3])
def test_hodge_diagonal_closed_form_and_double_star(n):
    """Diagonal metric, rectilinear axes: ⋆(I→J) = sign(I‖J)·Π_J √g h / Π_I √g h,
    and ⋆(J→I)·⋆(I→J) = (−1)^{k(N−k)}."""
    rng = np.random.default_rng(11 + n)
    ext = [4, 5, 3][:n]
    axes = [np.cumsum(np.concatenate([[0.0], rng.uniform(0.5, 1.5, e - 1)])) for e in ext]
    g = rng.uniform(0.5, 2.0, size=(int(np.prod(ext)), n))
    geom = orc.Geometry(ext, axes=axes, metric_kind=orc.METRIC_DIAGONAL, metric=g)
    for node in [[0] * n, [e - 1 for e in ext], [1] * n]:
        lin = int(np.ravel_multi_index(node, ext))
        s = []
        for a in range(n):
            i = node[a]
            h = axes[a][i + 1] - axes[a][i] if i + 1 < ext[a] else axes[a][i] - axes[a][i - 1]
            s.append(math.sqrt(g[lin, a]) * h)
        for k in range(n + 1):
            for src in combinations(range(n), k):
                sm = sum(1 << a for a in src)
                tm = ((1 << n) - 1) & ~sm
                ids = list(src) + [a for a in range(n) if not tm >> a & 1 and a not in src] + \
                    [a for a in range(n) if tm >> a & 1]
                perm = list(src) + [a for a in range(n) if tm >> a & 1]
                inv = sum(1 for i in range(n) for j in range(i + 1, n) if perm[i] > perm[j])
                expect = (-1) ** inv * np.prod([s[a] for a in range(n) if tm >> a & 1]) / np.prod(
                    [s[a] for a in src])
                got = orc.hodge_coefficient(geom, node, sm, tm)
                assert math.isclose(got, expect, rel_tol=1e-14)
                back = orc.hodge_coefficient(geom, node, tm, sm)
                assert math.isclose(got * back, (-1) ** (k * (n - k)), rel_tol=1e-14)


def test_codifferential_reference_closed_form():
    """tests/exterior/codifferential.cpp:40-117: 31×31 points on [-5,5]²,
    ω = x·dx dx + y·dy dy with dx = dy = 10/31, identity metric:
    δω = −2·30/31 on interior nodes at 1e-12."""
    n = 31
    c = np.linspace(-5.0, 5.0, n)
    d = 10.0 / 31.0
    X, Y = np.meshgrid(c, c, indexing="ij")
    form = np.stack([X * d, Y * d], axis=-1)
    metric = np.tile([1.0, 0.0, 1.0], (n * n, 1))
    geom = orc.Geometry([n, n], axes=[c, c], metric_kind=orc.METRIC_SYMMETRIC, metric=metric)
    out = orc.codifferential(geom, 1, form, [n, n])
    inner = out[1:-1, 1:-1, 0]
    assert np.max(np.abs(inner - (-2.0 * 30.0 / 31.0))) <= 1e-12


def gate_geometry(n, npts):
    c = np.linspace(-5.0, 5.0, npts)
    return orc.Geometry([npts] * n, axes=[c] * n), c


def test_laplacian_gate_1d_0form():
    """tests/exterior/laplacian.cpp:380-430"""
    geom, c = gate_geometry(1, 1000)
    R, a = 2.0, 0.5
    r = np.abs(c)
    f = np.where(r <= R, -a * r * r, -a * R * (2 * r - R))[:, None]
    lap = orc.laplacian(geom, 0, f, [999])[:, 0]
    xs = c[:999]
    sel = slice(1, 998)
    v, x = lap[sel], xs[sel]
    assert np.all(np.abs(v[(x < -1.2 * R) | (x > 1.2 * R)]) <= 1e-2)
    assert np.all(np.abs(v[(x > -0.8 * R) & (x < 0.8 * R)] - 1.0) <= 1e-2)


def test_laplacian_gate_1d_1form():
    """tests/exterior/laplacian.cpp:434-483"""
    geom, c = gate_geometry(1, 1000)
    R, a = 2.0, 0.5
    r = np.abs(c)
    f = np.where(r <= R, -a * r * r, -a * R * (2 * r - R))[:, None]
    lap = orc.laplacian(geom, 1, f, [999])[:, 0]
    x = c[:999][1:998]
    v = lap[1:998]
    assert np.all(np.abs(v[(x < -1.2 * R) | (x > 1.2 * R)]) <= 1e-2)
    assert np.all(np.abs(v[(x > -0.8 * R) & (x < 0.8 * R)] - 1.0) <= 1e-2)


def test_laplacian_gate_2d_0form():
    """tests/exterior/laplacian.cpp:487-556"""
    geom, c = gate_geometry(2, 31)
    R, a = 2.0, 0.25
    X, Y = np.meshgrid(c, c, indexing="ij")
    r = np.sqrt(X * X + Y * Y)
    with np.errstate(divide="ignore"):
        f = np.where(r <= R, -a * r * r, a * R * R * (2 * np.log(R / np.maximum(r, 1e-300)) - 1))
    lap = orc.laplacian(geom, 0, f[..., None], [30, 30])[..., 0]
    v = lap[1:29, 15]
    x = c[1:29]
    assert np.all(np.abs(v[(x < -1.2 * R) | (x > 1.2 * R)]) <= 0.5)
    assert np.all(np.abs(v[(x > -0.8 * R) & (x < 0.8 * R)] - 1.0) <= 0.5)


def test_laplacian_gate_2d_1form():
    """tests/exterior/laplacian.cpp:558-628"""
    geom, c = gate_geometry(2, 31)
    R = 2.0
    X, Y = np.meshgrid(c, c, indexing="ij")
    r = np.sqrt(X * X + Y * Y)
    with np.errstate(divide="ignore", invalid="ignore"):
        fac = np.where(r <= R, r / 3.0 - R / 2.0, -R ** 3 / (6.0 * r * r))
    form = np.stack([Y * fac, -X * fac], axis=-1)
    lap = orc.laplacian(geom, 1, form, [30, 30])
    v = lap[1:29, 15, 1]
    x = c[1:29]
    assert np.all(np.abs(v[(x < -1.2 * R) | (x > 1.2 * R)]) <= 0.5)
    assert np.all(np.abs(v[(x > -0.8 * R) & (x < -0.2 * R)] + 1.0) <= 0.5)
    assert np.all(np.abs(v[(x > 0.2 * R) & (x < 0.8 * R)] - 1.0) <= 0.5)


def swirl_3d(c, two_form):
    R, a = 2.0, 0.2
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    r = np.sqrt(X * X + Y * Y + Z * Z)
    th = np.arctan2(Y, X)
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.where(r <= R, a * r * r, a * R ** 4 / (r * r))
    if not two_form:
        return np.stack([amp * np.sin(th), -amp * np.cos(th), np.zeros_like(r)], axis=-1)
    # components XY, XZ, YZ
    return np.stack([np.zeros_like(r), amp * np.cos(th), amp * np.sin(th)], axis=-1)


def test_laplacian_gate_3d_1form():
    """tests/exterior/laplacian.cpp:745-835"""
    geom, c = gate_geometry(3, 21)
    lap = orc.laplacian(geom, 1, swirl_3d(c, False), [20, 20, 20])
    v = lap[1:19, 10, 10, 1]
    x = c[1:19]
    R = 2.0
    assert np.all(np.abs(v[(x < -1.3 * R) | (x > 1.3 * R)]) <= 0.5)
    assert np.all(np.abs(v[(x > -0.7 * R) & (x < -0.3 * R)] + 1.0) <= 0.5)
    assert np.all(np.abs(v[(x > 0.3 * R) & (x < 0.7 * R)] - 1.0) <= 0.5)


def test_laplacian_gate_3d_2form():
    """tests/exterior/laplacian.cpp:837-923"""
    geom, c = gate_geometry(3, 21)
    lap = orc.laplacian(geom, 2, swirl_3d(c, True), [20, 20, 20])
    v = lap[1:19, 10, 10, 1]
    x = c[1:19]
    R = 2.0
    assert np.all(np.abs(v[(x < -1.3 * R) | (x > 1.3 * R)]) <= 0.5)
    assert np.all(np.abs(v[(x > -0.7 * R) & (x < -0.3 * R)] - 1.0) <= 0.5)
    assert np.all(np.abs(v[(x > 0.3 * R) & (x < 0.7 * R)] + 1.0) <= 0.5)


def test_laplacian_0form_uniform_is_five_point_stencil():
    """Identity metric, uniform spacing: Δf = (4f − Σ neighbours)/h² inside,
    and the i−1 / j−1 terms drop at the first node of each axis."""
    rng = np.random.default_rng(5)
    n, h = 9, 0.25
    c = np.arange(n) * h
    geom = orc.Geometry([n, n], axes=[c, c])
    f = rng.standard_normal((n, n))
    lap = orc.laplacian(geom, 0, f[..., None], [n - 1, n - 1])[..., 0]
    fp = np.pad(f, 1)
    ref = np.zeros((n - 1, n - 1))
    for i in range(n - 1):
        for j in range(n - 1):
            I, J = i + 1, j + 1
            s = (fp[I + 1, J] - fp[I, J]) + (fp[I, J + 1] - fp[I, J])
            if i > 0:
                s -= fp[I, J] - fp[I - 1, J]
            if j > 0:
                s -= fp[I, J] - fp[I, J - 1]
            ref[i, j] = -s / (h * h)
    np.testing.assert_allclose(lap, ref, rtol=1e-13, atol=1e-12)


def test_laplacian_1form_uniform_is_vector_laplacian_inside():
    """Identity metric, uniform spacing: on nodes two steps away from every
    boundary, δd + dδ on a 1-form is the component-wise 7-point −∇²."""
    rng = np.random.default_rng(6)
    n, h = 8, 0.5
    c = np.arange(n) * h
    geom = orc.Geometry([n] * 3, axes=[c] * 3)
    w = rng.standard_normal((n, n, n, 3))
    lap = orc.laplacian(geom, 1, w, [n - 1] * 3)
    s = slice(2, n - 2)
    ref = 6 * w.copy()
    for ax in range(3):
        ref -= np.roll(w, 1, axis=ax) + np.roll(w, -1, axis=ax)
    ref /= h * h
    np.testing.assert_allclose(lap[s, s, s], ref[s, s, s], rtol=1e-12, atol=1e-11)


def test_laplacian_of_constant_0form_vanishes():
    rng = np.random.default_rng(8)
    axes = [np.cumsum(rng.uniform(0.5, 2.0, 7)) for _ in range(3)]
    g = rng.uniform(0.5, 2.0, size=(7 ** 3, 3))
    geom = orc.Geometry([7] * 3, axes=axes, metric_kind=orc.METRIC_DIAGONAL, metric=g)
    lap = orc.laplacian(geom, 0, np.full((7, 7, 7, 1), 3.25), [6] * 3)
    assert np.max(np.abs(lap)) <= 1e-12


def test_weighted_0form_laplacian_is_symmetric():
    """⋆0·Δ on sequences vanishing on the boundary is d0ᵀ⋆1d0 (symmetric)."""
    rng = np.random.default_rng(9)
    n = 7
    axes = [np.cumsum(rng.uniform(0.5, 2.0, n)) for _ in range(2)]
    geom = orc.Geometry([n, n], axes=axes)
    inner = [(i, j) for i in range(1, n - 1) for j in range(1, n - 1)]
    M = np.zeros((len(inner), len(inner)))
    for col, (i, j) in enumerate(inner):
        f = np.zeros((n, n, 1))
        f[i, j, 0] = 1.0
        lap = orc.laplacian(geom, 0, f, [n - 1, n - 1])[..., 0]
        for row, (p, q) in enumerate(inner):
            hx = axes[0][p + 1] - axes[0][p]
            hy = axes[1][q + 1] - axes[1][q]
            M[row, col] = hx * hy * lap[p, q]
    np.testing.assert_allclose(M, M.T, rtol=1e-12, atol=1e-12)


def test_reference_structured_cpu_baseline_matches_oracle():
    """The staged six-pass CPU baseline computes the same Δ as the oracle."""
    rng = np.random.default_rng(12)
    ext = [9, 8, 10]
    axes = [np.cumsum(np.concatenate([[0.0], rng.uniform(0.5, 1.5, e - 1)])) for e in ext]
    g = rng.uniform(0.5, 2.0, size=(int(np.prod(ext)), 3))
    geom = orc.Geometry(ext, axes=axes, metric_kind=orc.METRIC_DIAGONAL, metric=g)
    w = rng.standard_normal(tuple(ext) + (3,))
    ref = orc.laplacian(geom, 1, w, [e - 1 for e in ext])
    got = orc.StagedLaplacian3D1(ext, axes, g).run(w, [e - 1 for e in ext])
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
