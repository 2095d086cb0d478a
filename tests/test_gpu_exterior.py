"""GPU parity of the sil::exterior operators (through the C ABI) against the
CPU oracle: bit-exact for the coboundary on integer-valued cochains, ≤1e-12
relative (‖·‖∞ norm-wise) in FP64 for ⋆, δ and Δ."""
import json
import os
from itertools import combinations
from math import comb
from pathlib import Path

import numpy as np
import pytest
import torch

import oracle_lib as orc

pytestmark = pytest.mark.gpu

TOL = 1e-12  # relative, norm-wise ‖y − y_ref‖∞ / ‖y_ref‖∞


@pytest.fixture(scope="module")
def ext():
    import similie_amd.exterior as E

    return E


def cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def rel_err(got, ref):
    got = got.detach().cpu().numpy() if isinstance(got, torch.Tensor) else got
    den = np.max(np.abs(ref))
    return np.max(np.abs(got - ref)) / (den if den > 0 else 1.0)


def harness_input(ndim, npts, degree):
    nc = comb(ndim, degree)
    arr = np.zeros((npts,) * ndim + (nc,))
    for i in range(nc):
        arr[..., i] = i + 1.0
        arr[(1,) * ndim + (i,)] = i + 2.0
    return arr


CASES = json.loads((Path(__file__).with_name("golden") / "derivative_known_answers.json").read_text())["cases"]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['test']}-{'first' if c['index_first'] else 'last'}")
def test_coboundary_reference_known_answers(ext, case):
    n, npts, k = case["ndim"], case["points_per_axis"], case["in_degree"]
    arr = harness_input(n, npts, k)
    oext = [npts if case["out_degree"] == 1 else npts - 1] * n
    nout = comb(n, k + 1)
    if case["index_first"]:
        t = cuda(np.moveaxis(arr, -1, 0))
        out = torch.full((nout,) + tuple(oext), np.nan, dtype=torch.float64, device="cuda")
        ext.coboundary(out, t, k, comp_dim=0)
        got = np.moveaxis(out.cpu().numpy(), 0, -1)
    else:
        t = cuda(arr)
        out = torch.full(tuple(oext) + (nout,), np.nan, dtype=torch.float64, device="cuda")
        ext.coboundary(out, t, k)
        got = out.cpu().numpy()
    ref = orc.coboundary(n, k, [npts] * n, arr, oext)
    np.testing.assert_array_equal(got, ref)
    for e in case["expects"]:
        comp = orc.combo_index(n, k + 1, e["axes_mask"])
        assert got[tuple(e["node"]) + (comp,)] == e["value"]


def test_coboundary_with_spectator_dimensions(ext):
    """tests/exterior/derivative.cpp:889-990: layout
    [Antisym<Mu2>][Spect][X][IndexSpect (symmetric 4×4: 10)][Y]."""
    N = 3
    t = torch.zeros((2, N, N, 10, N), dtype=torch.float64, device="cuda")
    for i in range(2):
        t[i] = i + 1.0
        t[i, :, 1, :, 1] = i + 2.0
    out = torch.full((1, 3, 2, 10, 2), np.nan, dtype=torch.float64, device="cuda")
    ext.coboundary(out, t, 1, grid_dims=[2, 4], comp_dim=0, batch_dims=[1, 3])
    o = out.cpu().numpy()
    assert np.all(o[0, :, 0, :, 0] == 0.0)
    assert np.all(o[0, :, 1, :, 0] == -1.0)
    assert np.all(o[0, :, 0, :, 1] == 1.0)
    assert np.all(o[0, :, 1, :, 1] == 0.0)


@pytest.mark.parametrize("n,k", [(1, 0), (2, 0), (2, 1), (3, 0), (3, 1), (3, 2), (4, 0), (4, 1), (4, 2), (4, 3)])
def test_coboundary_integer_cochains_bit_exact(ext, n, k):
    rng = np.random.default_rng(100 + 10 * n + k)
    e = [7, 5, 6, 4][:n]
    arr = rng.integers(-1000, 1001, size=tuple(e) + (comb(n, k),)).astype(np.float64)
    for oext in (e, [x - 1 for x in e], [x + 1 for x in e]):
        out = torch.empty(tuple(oext) + (comb(n, k + 1),), dtype=torch.float64, device="cuda")
        ext.coboundary(out, cuda(arr), k)
        np.testing.assert_array_equal(out.cpu().numpy(), orc.coboundary(n, k, e, arr, oext))
        tout = torch.empty_like(out)
        ext.transposed_coboundary(tout, cuda(arr), k)
        np.testing.assert_array_equal(tout.cpu().numpy(), orc.transposed_coboundary(n, k, e, arr, oext))


def test_coboundary_3d_0form_config_size_bit_exact(ext):
    """256³ integer 0-cochain in [-1000, 1000] → 1-form, bit-exact; checked
    against the closed form on the whole grid (the oracle runs on crops)."""
    g = torch.Generator(device="cuda").manual_seed(7)
    f = torch.randint(-1000, 1001, (256, 256, 256, 1), generator=g, device="cuda").to(torch.float64)
    out = torch.empty((256, 256, 256, 3), dtype=torch.float64, device="cuda")
    ext.coboundary(out, f, 0)
    ff = f[..., 0]
    for a in range(3):
        fwd = torch.cat([ff.narrow(a, 1, 255), ff.narrow(a, 255, 1)], dim=a)
        assert torch.equal(out[..., a], fwd - ff)
    crop = f[100:108, 40:48, 248:256].cpu().numpy()
    ref = orc.coboundary(3, 0, [8, 8, 8], crop, [8, 8, 8])
    got = out[100:108, 40:48, 248:256].cpu().numpy()
    np.testing.assert_array_equal(got[:7, :7, :], ref[:7, :7, :])


def geometry(n, extents, kind, rng):
    axes = [np.cumsum(np.concatenate([[0.0], rng.uniform(0.5, 1.5, e - 1)])) for e in extents]
    nodes = int(np.prod(extents))
    if kind == "identity":
        return axes, None, orc.METRIC_IDENTITY, None
    if kind == "diagonal":
        m = rng.uniform(0.5, 2.0, size=(nodes, n))
        return axes, None, orc.METRIC_DIAGONAL, m
    sym = []
    for _ in range(nodes):
        a = rng.standard_normal((n, n))
        s = a @ a.T + n * np.eye(n)
        sym.append([s[i, j] for i in range(n) for j in range(i, n)])
    return axes, None, orc.METRIC_SYMMETRIC, np.array(sym)


def make_pair(E, n, extents, kind, rng, nodal=False):
    axes, _, mk, metric = geometry(n, extents, kind, rng)
    if nodal:
        base = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1)
        pos = base + 0.1 * rng.standard_normal(base.shape)
        og = orc.Geometry(extents, nodal=pos.reshape(-1, n), metric_kind=mk, metric=metric)
        gg = E.Grid(extents, nodal_positions=cuda(pos), metric=cuda(metric) if metric is not None else None,
                    metric_kind=mk)
    else:
        og = orc.Geometry(extents, axes=axes, metric_kind=mk, metric=metric)
        gg = E.Grid(extents, axis_coords=axes, metric=cuda(metric) if metric is not None else None, metric_kind=mk)
    return og, gg


@pytest.mark.parametrize("kind", ["identity", "diagonal", "symmetric"])
@pytest.mark.parametrize("n", [2, 3])
def test_hodge_star_fill_and_apply(ext, n, kind):
    rng = np.random.default_rng(3 + n)
    extents = [6, 5, 4][:n]
    for nodal in (False, True):
        og, gg = make_pair(ext, n, extents, kind, rng, nodal)
        for k in range(n + 1):
            ref = orc.fill_hodge_star(og, k)
            got = ext.fill_discrete_hodge_star(gg, k)
            assert rel_err(got, ref) <= TOL
            w = rng.standard_normal(tuple(extents) + (comb(n, k),))
            out = torch.empty(tuple(extents) + (comb(n, n - k),), dtype=torch.float64, device="cuda")
            ext.hodge_star(gg, k, out, cuda(w))
            assert rel_err(out, orc.hodge_star(og, k, w)) <= TOL


@pytest.mark.parametrize("kind", ["identity", "diagonal", "symmetric"])
@pytest.mark.parametrize("n", [1, 2, 3])
def test_codifferential_matches_oracle(ext, n, kind):
    rng = np.random.default_rng(30 + n)
    extents = [7, 6, 5][:n]
    og, gg = make_pair(ext, n, extents, kind, rng)
    for k in range(1, n + 1):
        w = rng.standard_normal(tuple(extents) + (comb(n, k),))
        for oext in (extents, [e - 1 for e in extents]):
            out = torch.empty(tuple(oext) + (comb(n, k - 1),), dtype=torch.float64, device="cuda")
            ext.codifferential(gg, k, out, cuda(w))
            assert rel_err(out, orc.codifferential(og, k, w, oext)) <= TOL


def test_codifferential_reference_closed_form(ext):
    n = 31
    c = np.linspace(-5.0, 5.0, n)
    d = 10.0 / 31.0
    X, Y = np.meshgrid(c, c, indexing="ij")
    form = np.stack([X * d, Y * d], axis=-1)
    metric = np.tile([1.0, 0.0, 1.0], (n * n, 1))
    gg = ext.Grid([n, n], axis_coords=[c, c], metric=cuda(metric), metric_kind=2)
    out = torch.empty((n, n, 1), dtype=torch.float64, device="cuda")
    ext.codifferential(gg, 1, out, cuda(form))
    inner = out.cpu().numpy()[1:-1, 1:-1, 0]
    assert np.max(np.abs(inner + 2.0 * 30.0 / 31.0)) <= 1e-12


@pytest.mark.parametrize("kind", ["identity", "diagonal", "symmetric"])
@pytest.mark.parametrize("n,k", [(1, 0), (1, 1), (2, 0), (2, 1), (2, 2), (3, 0), (3, 1), (3, 2), (3, 3)])
def test_laplacian_matches_oracle(ext, n, k, kind):
    rng = np.random.default_rng(50 + 10 * n + k)
    extents = [9, 7, 6][:n]
    og, gg = make_pair(ext, n, extents, kind, rng)
    w = rng.standard_normal(tuple(extents) + (comb(n, k),))
    for oext in ([e - 1 for e in extents], extents):
        out = torch.empty(tuple(oext) + (comb(n, k),), dtype=torch.float64, device="cuda")
        ext.laplacian(gg, k, out, cuda(w))
        assert rel_err(out, orc.laplacian(og, k, w, oext)) <= TOL


TILES = ["0,8,32", "1,4,3", "1,8,32", "1,8,5", "1,16,7", "1,8,1", "1,12,32", "1,12,5",
         "2,8,5", "2,16,7", "2,12,32", "2,8,1", "2,16,64", "2,4,3,256", "2,4,7,128", "2,8,32,128", "2,4,2,256",
         "3,8,5", "3,12,32", "3,16,7", "3,8,1", "3,16,5,1", "3,16,32,3", "3,16,1,3"]


@pytest.mark.parametrize("tile", TILES)
@pytest.mark.parametrize("kind", ["identity", "diagonal"])
def test_fused_3d_1form_laplacian_tiles(ext, kind, tile, monkeypatch):
    """Fused kernels (direct and x-marching with several tiles) on a
    stretched grid whose extents are not multiples of any tile."""
    monkeypatch.setenv("SIL_LAP3D_TILE", tile)
    rng = np.random.default_rng(77)
    extents = [19, 23, 71]
    og, gg = make_pair(ext, 3, extents, kind, rng)
    w = rng.standard_normal(tuple(extents) + (3,))
    ref_small = orc.laplacian(og, 1, w, [e - 1 for e in extents])
    out = torch.full(tuple(e - 1 for e in extents) + (3,), np.nan, dtype=torch.float64, device="cuda")
    ext.laplacian(gg, 1, out, cuda(w))
    assert rel_err(out, ref_small) <= TOL
    full = torch.full(tuple(extents) + (3,), np.nan, dtype=torch.float64, device="cuda")
    ext.laplacian(gg, 1, full, cuda(w))
    assert rel_err(full, orc.laplacian(og, 1, w, extents)) <= TOL


@pytest.mark.parametrize("kind", ["identity", "diagonal"])
def test_fused_2d_0form_laplacian(ext, kind):
    rng = np.random.default_rng(88)
    extents = [37, 41]
    og, gg = make_pair(ext, 2, extents, kind, rng)
    f = rng.standard_normal(tuple(extents) + (1,))
    for oext in ([36, 40], extents):
        out = torch.empty(tuple(oext) + (1,), dtype=torch.float64, device="cuda")
        ext.laplacian(gg, 0, out, cuda(f))
        assert rel_err(out, orc.laplacian(og, 0, f, oext)) <= TOL


def test_staged_laplacian_reuse(ext):
    rng = np.random.default_rng(9)
    extents = [12, 10, 14]
    og, gg = make_pair(ext, 3, extents, "diagonal", rng)
    out = torch.empty((11, 9, 13, 3), dtype=torch.float64, device="cuda")
    w0 = cuda(rng.standard_normal(tuple(extents) + (3,)))
    staged = ext.make_staged_laplacian(gg, 1, out, w0)
    for s in range(3):
        w = rng.standard_normal(tuple(extents) + (3,))
        staged.run(out, cuda(w))
        assert rel_err(out, orc.laplacian(og, 1, w, [11, 9, 13])) <= TOL


def test_staged_laplacian_autotune(ext):
    """the autotuner times the fused kernel shapes on the staged operator's
    arrays, keeps one, leaves its result in out, and later runs use it"""
    rng = np.random.default_rng(13)
    extents = [40, 36, 70]
    og, gg = make_pair(ext, 3, extents, "diagonal", rng)
    oext = [e - 1 for e in extents]
    out = torch.empty(tuple(oext) + (3,), dtype=torch.float64, device="cuda")
    w = rng.standard_normal(tuple(extents) + (3,))
    wd = cuda(w)
    staged = ext.make_staged_laplacian(gg, 1, out, wd)
    tile = staged.autotune(out, wd, repeats=2)
    print(f"[laplacian autotune] {tile}", flush=True)
    assert tile["variant"] in ("stage", "pipe") and tile["microseconds"] > 0.0
    ref = orc.laplacian(og, 1, w, oext)
    assert rel_err(out, ref) <= TOL
    out.zero_()
    staged.run(out, wd)
    assert rel_err(out, ref) <= TOL


@pytest.mark.parametrize("n,k", [(3, 1), (3, 2), (3, 3), (2, 1), (2, 2)])
def test_staged_codifferential_reuse(ext, n, k):
    """StagedCodifferential (codifferential.hpp:321-508): one staged handle
    run on several forms gives the one-shot codifferential bit for bit and
    the oracle's within the DEC tolerance"""
    from math import comb

    rng = np.random.default_rng(20 + 3 * n + k)
    extents = [9, 7, 8][:n]
    og, gg = make_pair(ext, n, extents, "diagonal", rng)
    oext = [e - 1 for e in extents]
    out = torch.empty(tuple(oext) + (comb(n, k - 1),), dtype=torch.float64, device="cuda")
    one = torch.empty_like(out)
    w0 = cuda(rng.standard_normal(tuple(extents) + (comb(n, k),)))
    staged = ext.make_staged_codifferential(gg, k, out, w0)
    for _ in range(3):
        w = rng.standard_normal(tuple(extents) + (comb(n, k),))
        staged.run(out, cuda(w))
        ext.codifferential(gg, k, one, cuda(w))
        assert torch.equal(out, one)
        assert rel_err(out, orc.codifferential(og, k, w, oext)) <= TOL
    with pytest.raises(ValueError):
        staged.run(out[:-1], cuda(w))


def config2_grid(ext, n=256, seed=2):
    """Geometric spacing ratio 1.01 per cell on [0, 1] and diagonal metric
    g_dd = 1 + x_d² (the 256³ throughput workload)."""
    h = 1.01 ** np.arange(n - 1)
    c = np.concatenate([[0.0], np.cumsum(h)]) / h.sum()
    X = torch.from_numpy(c).cuda()
    g = torch.empty((n, n, n, 3), dtype=torch.float64, device="cuda")
    g[..., 0] = (1 + X * X).view(n, 1, 1)
    g[..., 1] = (1 + X * X).view(1, n, 1)
    g[..., 2] = (1 + X * X).view(1, 1, n)
    grid = ext.Grid([n] * 3, axis_coords=[c] * 3, metric=g)
    return grid, c, g


@pytest.mark.parametrize("tile", [None, "2,12,32", "3,16,32"])
def test_fused_3d_1form_laplacian_full_size(ext, monkeypatch, tile):
    """At 256³: the x-marching kernels (the default and each variant) against
    the direct kernel everywhere, and against the oracle on crops whose local
    result is exactly the full grid's (interior crops, and crops touching the
    upper faces)."""
    grid, c, g = config2_grid(ext)
    gen = torch.Generator(device="cuda").manual_seed(11)
    w = torch.randn((256, 256, 256, 3), generator=gen, dtype=torch.float64, device="cuda")
    out = torch.full((255, 255, 255, 3), np.nan, dtype=torch.float64, device="cuda")
    if tile is not None:
        monkeypatch.setenv("SIL_LAP3D_TILE", tile)
    ext.laplacian(grid, 1, out, w)
    monkeypatch.setenv("SIL_LAP3D_TILE", "0,8,32")
    out_d = torch.empty_like(out)
    ext.laplacian(grid, 1, out_d, w)
    err = (out - out_d).abs().max().item() / out_d.abs().max().item()
    assert err <= TOL
    gm = g.cpu().numpy()
    wn = w.cpu().numpy()
    for lo in ([0, 0, 0], [120, 37, 200], [248, 248, 248]):
        sl = tuple(slice(l, l + 8) for l in lo)
        og = orc.Geometry([8] * 3, axes=[c[l:l + 8] for l in lo], metric_kind=orc.METRIC_DIAGONAL,
                          metric=gm[sl].reshape(-1, 3))
        ref = orc.laplacian(og, 1, wn[sl], [7] * 3)
        got = out[tuple(slice(l, l + 7) for l in lo)].cpu().numpy()
        # nodes at crop index 0 see the crop's zero extension unless the crop starts at 0;
        # nodes at crop index 6 need δω at 7, whose backward edges match only at the grid end
        inner = tuple(slice(0 if l == 0 else 1, 7 if l + 8 == 256 else 6) for l in lo)
        assert rel_err(got[inner], ref[inner]) <= TOL
