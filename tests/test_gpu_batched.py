"""GPU checks of the batched persistent Poisson solver: every problem
reaches an FP64-verified 1e-10 relative residual (recomputed here with the
CPU oracle's operator), for every workgroup size the autotuner chooses from."""
import numpy as np
import pytest
import torch

import oracle_lib as orc

pytestmark = pytest.mark.gpu


def problems(seed, count, lo=8, hi=70):
    rng = np.random.default_rng(seed)
    sides = np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(int)
    kap, rhs = [], []
    for n in sides:
        kap.append(np.exp(rng.standard_normal((n, n))))
        b = np.zeros((n, n))
        b[1:-1, 1:-1] = rng.standard_normal((n - 2, n - 2))
        rhs.append(b)
    return sides, kap, rhs


@pytest.mark.parametrize("block", [256, 512, 1024])
def test_batched_solve_reaches_fp64_verified_tolerance(block):
    from similie_amd.batched import BatchedPoisson2D

    sides, kap, rhs = problems(3, 24)
    kappa = torch.from_numpy(np.concatenate([k.ravel() for k in kap])).cuda()
    b = torch.from_numpy(np.concatenate([r.ravel() for r in rhs])).cuda()
    solver = BatchedPoisson2D(sides, kappa)
    solver.set_block(block)
    x, results, secs = solver.solve(b)
    xh = x.cpu().numpy()
    for p, n in enumerate(sides):
        assert results[p]["converged"] == 1, (p, n, results[p])
        assert results[p]["true_relative_residual"] <= 1e-10
        o = solver.offset(p)
        xp = xh[o:o + n * n].reshape(n, n)
        ones = np.ones(n)
        r = rhs[p] - orc.poisson2d_apply(ones, ones, kap[p], xp)
        assert np.linalg.norm(r) <= 1.0001e-10 * np.linalg.norm(rhs[p])


def test_autotune_picks_a_candidate_and_solves():
    from similie_amd.batched import BatchedPoisson2D

    sides, kap, rhs = problems(4, 40, 16, 120)
    kappa = torch.from_numpy(np.concatenate([k.ravel() for k in kap])).cuda()
    b = torch.from_numpy(np.concatenate([r.ravel() for r in rhs])).cuda()
    solver = BatchedPoisson2D(sides, kappa)
    block, secs = solver.autotune(b, probe_iterations=10)
    assert block in (256, 512, 1024) and all(v > 0 for v in secs.values())
    x, results, _ = solver.solve(b)
    assert all(r["converged"] == 1 and r["true_relative_residual"] <= 1e-10 for r in results)


@pytest.mark.parametrize("wide_side", ["3", "40", "100000"])
def test_wide_path_reaches_fp64_verified_tolerance(wide_side, monkeypatch):
    """Problems at or above the wide side are spread over many workgroups
    (kernel boundaries as the only barriers); the rest stay one workgroup
    each.  Every split — all wide, mixed, none wide — solves every problem."""
    from similie_amd.batched import BatchedPoisson2D

    monkeypatch.setenv("SIL_BATCH_WIDE_SIDE", wide_side)
    sides, kap, rhs = problems(5, 18, 8, 90)
    kappa = torch.from_numpy(np.concatenate([k.ravel() for k in kap])).cuda()
    b = torch.from_numpy(np.concatenate([r.ravel() for r in rhs])).cuda()
    solver = BatchedPoisson2D(sides, kappa)
    x, results, _ = solver.solve(b)
    xh = x.cpu().numpy()
    for p, n in enumerate(sides):
        assert results[p]["converged"] == 1, (p, n, results[p])
        assert results[p]["true_relative_residual"] <= 1e-10
        o = solver.offset(p)
        xp = xh[o:o + n * n].reshape(n, n)
        ones = np.ones(n)
        r = rhs[p] - orc.poisson2d_apply(ones, ones, kap[p], xp)
        assert np.linalg.norm(r) <= 1.0001e-10 * np.linalg.norm(rhs[p])
    # a second solve on the same handle reuses the captured graphs
    x2, results2, _ = solver.solve(b)
    assert [r["iterations"] for r in results2] == [r["iterations"] for r in results]


def test_shape_autotune_keeps_every_problem_solved():
    """The shape autotuner re-partitions the batch (wide-path threshold,
    workgroups per wide problem, workgroup size) between timed probes; the
    shape it leaves behind still solves every problem to the verified
    tolerance, and a second solve repeats the first bit for bit."""
    from similie_amd.batched import BatchedPoisson2D

    sides, kap, rhs = problems(6, 30, 8, 260)
    kappa = torch.from_numpy(np.concatenate([k.ravel() for k in kap])).cuda()
    b = torch.from_numpy(np.concatenate([r.ravel() for r in rhs])).cuda()
    solver = BatchedPoisson2D(sides, kappa)
    shape = solver.autotune_shape(b, probe_iterations=8)
    assert shape["block"] in (256, 512, 1024) and shape["probe_seconds"] > 0
    assert shape["wide_side"] in (96, 128, 192, 256, 384, int(max(sides)) + 1)
    assert shape["wide_problems"] == sum(int(n) >= shape["wide_side"] for n in sides)
    assert shape["wide_workgroups"] >= 1
    x, results, _ = solver.solve(b)
    xh = x.cpu().numpy()
    for p, n in enumerate(sides):
        assert results[p]["converged"] == 1, (p, n, results[p])
        assert results[p]["true_relative_residual"] <= 1e-10
        o = solver.offset(p)
        xp = xh[o:o + n * n].reshape(n, n)
        ones = np.ones(n)
        r = rhs[p] - orc.poisson2d_apply(ones, ones, kap[p], xp)
        assert np.linalg.norm(r) <= 1.0001e-10 * np.linalg.norm(rhs[p])
    x2, results2, _ = solver.solve(b)
    assert torch.equal(x, x2) and [r["iterations"] for r in results2] == [r["iterations"] for r in results]
