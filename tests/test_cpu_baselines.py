"""The CPU baselines bench.py times (oracle/cpu_baseline.c, oracle/cpu_fast.c)
compute what the oracle computes: the fused one-sweep Laplacian and the
staged reference-structured one against orc.laplacian, the OpenMP CGs
against the oracle's textbook CG and operators — in the literal build and in
the -O3 -march=native build bench.py makes."""
import numpy as np
import pytest

import oracle_lib as orc


@pytest.fixture(scope="module", params=["literal", "native"])
def base(request, tmp_path_factory):
    if request.param == "literal":
        return orc.CpuBaseline()
    path = orc.build_native(tmp_path_factory.mktemp("native"))
    assert path is not None, "the native CPU baseline build failed"
    return orc.CpuBaseline(path)


@pytest.mark.parametrize("ext, oext", [([9, 7, 11], [8, 6, 10]), ([6, 8, 5], [6, 8, 5]), ([12, 10, 9], [11, 9, 8]),
                                       ([3, 3, 3], [2, 2, 2])])
def test_laplacian_baselines_match_the_oracle(base, ext, oext):
    rng = np.random.default_rng(sum(ext))
    axes = [np.cumsum(np.concatenate([[0.0], rng.uniform(0.5, 1.5, e - 1)])) for e in ext]
    g = rng.uniform(0.5, 2.0, size=tuple(ext) + (3,))
    w = rng.standard_normal(tuple(ext) + (3,))
    ref = orc.laplacian(orc.Geometry(ext, axes=axes, metric_kind=orc.METRIC_DIAGONAL, metric=g.reshape(-1, 3)), 1,
                        w, oext)
    scale = np.max(np.abs(ref))
    fused = base.fused_lap3d1(ext, axes, g.reshape(-1, 3), w, oext)
    assert np.max(np.abs(fused - ref)) <= 1e-12 * scale
    staged = base.staged(ext, axes, g.reshape(-1, 3)).run(w, oext)
    assert np.max(np.abs(staged - ref)) <= 1e-12 * scale


def test_poisson_cg_baseline(base):
    rng = np.random.default_rng(3)
    nx, ny = 65, 57
    hx, hy = np.full(nx, 1 / 64), np.full(ny, 1 / 56)
    kappa = np.exp(rng.standard_normal((nx, ny)))
    b = np.zeros((nx, ny))
    b[1:-1, 1:-1] = rng.standard_normal((nx - 2, ny - 2))
    x, it, _, _ = base.poisson2d_cg(hx, hy, kappa, b, 1e-10, 5000, jacobi=False)
    xr, itr, _, _ = orc.poisson2d_cg(hx, hy, kappa, b, 1e-10, 5000)
    # the native build contracts multiply-adds: a few iterations either way
    assert abs(it - itr) <= 2 + 0.01 * itr
    assert np.max(np.abs(x - xr)) <= 1e-8 * np.max(np.abs(xr))
    xj, itj, _, _ = base.poisson2d_cg(hx, hy, kappa, b, 1e-10, 5000, jacobi=True)
    r = b - orc.poisson2d_apply(hx, hy, kappa, xj)
    assert np.linalg.norm(r) <= 1.5e-10 * np.linalg.norm(b) and itj < it
    _, fixed, _, _ = base.poisson2d_cg(hx, hy, kappa, b, 1e-10, -7)
    assert fixed == 7


@pytest.mark.parametrize("jacobi", [False, True])
def test_magnetostatics_cg_baseline(base, jacobi):
    rng = np.random.default_rng(5)
    shape = (12, 10, 14)
    coords = [np.sort(np.concatenate([[0.0], rng.uniform(0, 1, n - 2), [1.0]])) for n in shape]
    cells = (shape[0] - 1) * (shape[1] - 1) * (shape[2] - 1)
    cmu = np.where(rng.random(cells) < 0.3, 1000.0, 1.0) * 4e-7 * np.pi
    mu = orc.ms_nodal_mu(shape, cmu, 4e-7 * np.pi)
    b = orc.ms_rhs(coords, rng.standard_normal((cells, 3)))
    x, it, _, _ = base.ms3d_cg(coords, mu, 1e7, b, 1e-10, 20000, jacobi=jacobi)
    diag = orc.ms_diagonal(coords, mu, 1e7) if jacobi else None
    xr, itr, _, _ = orc.ms_cg(coords, mu, 1e7, b, 1e-10, 20000, diag=diag)
    # the same algorithm; on this ill-conditioned system (penalty 1e7, μ
    # ratio 1000) rounding in another summation order moves the count a little
    assert abs(it - itr) <= max(5, 0.05 * itr)
    r = b - orc.ms_apply(coords, mu, 1e7, x)
    assert np.linalg.norm(r) <= 1.5e-10 * np.linalg.norm(b)
    assert np.max(np.abs(x - xr)) <= 1e-6 * np.max(np.abs(xr))
