"""Throughput of the MI355X Hodge Laplacian on its headline workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-seconds S] [--no-cpu-baseline]

One step is one application of the 3D Hodge Laplacian δd + dδ to a 1-form
on 256³ nodes (geometric spacing, ratio 1.01 per cell on every axis of
[0, 1]; diagonal metric g_dd = 1 + x_d²; FP64), inputs resident in HBM.
Prints one JSON line.  `value` is algorithmic GB/s = (input 1-form + metric
+ output 1-form bytes) × steps / wall time of the timed steps.
"""
import argparse
import json
import math
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Laplacian GB/s (% HBM peak); time-to-1e-10 residual, learned-PCG vs plain CG"
HBM_PEAK_GBPS = 8000.0
N = 256


def algorithmic_bytes(n=N):
    # input 1-form (3 doubles/node) + diagonal metric (3/node) + output on (n-1)^3 nodes
    return 8 * (3 * n ** 3 + 3 * n ** 3 + 3 * (n - 1) ** 3)


def workload(torch, E, n=N):
    import numpy as np

    h = 1.01 ** np.arange(n - 1)
    c = np.concatenate([[0.0], np.cumsum(h)]) / h.sum()
    x = torch.from_numpy(c).cuda()
    g = torch.empty((n, n, n, 3), dtype=torch.float64, device="cuda")
    g[..., 0] = (1 + x * x).view(n, 1, 1)
    g[..., 1] = (1 + x * x).view(1, n, 1)
    g[..., 2] = (1 + x * x).view(1, 1, n)
    gen = torch.Generator(device="cuda").manual_seed(1234)
    w = torch.randn((n, n, n, 3), generator=gen, dtype=torch.float64, device="cuda")
    out = torch.empty((n - 1, n - 1, n - 1, 3), dtype=torch.float64, device="cuda")
    grid = E.Grid([n] * 3, axis_coords=[c] * 3, metric=g)
    return grid, c, g, w, out


def profiled_traffic(tile):
    """HBM bytes per launch of the kernel shape `tile` from the committed
    rocprofv3 --pmc summary (one line per shape the autotuner can choose:
    hbm_bytes_per_launch[variant,rows,xchunk,lanes_or_depth,nontemporal]);
    None when that shape has no measured line."""
    p = ROOT / "profiles" / "laplacian_3d_1form_256_pmc.md"
    if not p.exists():
        return None
    key = f"{tile['variant']},{tile['rows']},{tile['xchunk']},{tile['lanes_or_depth']},{int(tile['nontemporal'])}"
    m = re.search(r"hbm_bytes_per_launch\[" + re.escape(key) + r"\]:\s*([0-9.eE+]+)", p.read_text())
    return float(m.group(1)) if m else None


_CPU = {}


def cpu_lib():
    """The CPU baselines (oracle/cpu_baseline.c, oracle/cpu_fast.c) built on
    this host with -O3 -march=native -fopenmp, once per run; the literal
    oracle build when that fails.  (CpuBaseline, build flags)"""
    if "base" not in _CPU:
        import tempfile

        sys.path.insert(0, str(ROOT / "tests"))
        import oracle_lib as orc

        path = orc.build_native(tempfile.mkdtemp(prefix="sil_cpu_"))
        _CPU["base"] = orc.CpuBaseline(path)
        _CPU["build"] = "-O3 -march=native -fopenmp" if path else "-O2 -fopenmp (the native build failed)"
    return _CPU["base"], _CPU["build"]


def _median_time(fn, seconds, least=5, most=20):
    fn()  # warm-up
    times = []
    t_end = time.perf_counter() + seconds
    while len(times) < least or (len(times) < most and time.perf_counter() < t_end):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return sorted(times)[len(times) // 2], len(times)


def cpu_baseline(seconds):
    """The full 256³ workload on the host cores: the reference-structured
    staged Laplacian (90 ⋆ doubles per node filled once, six OpenMP passes
    per application, oracle/cpu_baseline.c) is the value; the fused
    one-sweep CPU Laplacian (oracle/cpu_fast.c) is reported beside it."""
    base, build = cpu_lib()
    n = N
    h = 1.01 ** np.arange(n - 1)
    c = np.concatenate([[0.0], np.cumsum(h)]) / h.sum()
    gx = 1 + c * c
    g = np.empty((n, n, n, 3))
    g[..., 0] = gx[:, None, None]
    g[..., 1] = gx[None, :, None]
    g[..., 2] = gx[None, None, :]
    g = g.reshape(-1, 3)
    w = np.random.default_rng(1234).standard_normal((n, n, n, 3))
    out = np.empty((n - 1, n - 1, n - 1, 3))
    oext = [n - 1] * 3
    t0 = time.perf_counter()
    st = base.staged([n] * 3, [c] * 3, g)
    setup = time.perf_counter() - t0
    t_staged, runs = _median_time(lambda: st.run(w, oext, out), seconds)
    del st
    t_fused, runs_f = _median_time(lambda: base.fused_lap3d1([n] * 3, [c] * 3, g, w, oext, out), seconds)
    B = algorithmic_bytes(n)
    return {
        "value": round(B / t_staged / 1e9, 3),
        "unit": "GB/s",
        "cores": base.threads(),
        "kind": "port",
        "sample": f"the full 256^3 workload: reference-structured staged Laplacian (90 star doubles per node "
                  f"filled once in {setup:.1f} s, six OpenMP passes), median of {runs} applications; built "
                  f"{build}",
        "fused_value": round(B / t_fused / 1e9, 3),
        "fused_sample": f"one-sweep CPU Laplacian (per-thread x-plane ranges, rolling plane buffers), median of "
                        f"{runs_f} applications",
    }


def _guard(name, fn):
    """Runs one solver measurement with a progress line on stderr; a failure
    is reported in the line instead of ending the benchmark."""
    t0 = time.perf_counter()
    print(f"bench: {name} ...", file=sys.stderr, flush=True)
    try:
        out = fn()
    except Exception as exc:  # noqa: BLE001 - reported, not swallowed
        out = {"error": f"{type(exc).__name__}: {exc}"}
    print(f"bench: {name} done in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    return out


def laplacian_2d_0form(torch):
    """Δ of a 0-form on the 256² Cartesian grid (spacing 1/256, Euclidean), N(0,1) values."""
    from similie_amd import exterior as E

    n = 256
    x = np.arange(n) / 256.0
    grid = E.Grid([n, n], axis_coords=[x, x])
    gen = torch.Generator(device="cuda").manual_seed(11)
    f = torch.randn((n, n, 1), generator=gen, dtype=torch.float64, device="cuda")
    out = torch.empty((n - 1, n - 1, 1), dtype=torch.float64, device="cuda")
    st = E.make_staged_laplacian(grid, 0, out, f)
    for _ in range(10):
        st.run(out, f)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(200):
        st.run(out, f)
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) / 200 * 1e3
    nbytes = 8 * (n * n + (n - 1) ** 2)
    return {"workload": "0-form Laplacian on 256^2, h = 1/256, Euclidean metric", "us_per_application": round(us, 2),
            "GBps": round(nbytes / (us * 1e-6) / 1e9, 1),
            "note": "1 MB in + out: the application is launch-latency bound, not HBM bound"}


def solve_poisson_bumps(torch, cpu=True):
    """256² Dirichlet Poisson, three Gaussian bumps, plain FP64 Jacobi CG to
    1e-10; the same solve on the host cores beside it."""
    from similie_amd import solvers as S
    from similie_amd.workloads import poisson_bumps

    x, y, b = poisson_bumps()
    op = S.Poisson2D(x, y)
    rhs = torch.from_numpy(b.reshape(-1)).cuda()
    st = S.StrongFormulationSolverSettings(max_iterations=20000, relative_tolerance=1e-10)
    S.minimize_strong_formulation_residual(op, rhs, settings=st)  # warm-up (graph capture)
    _, d = S.minimize_strong_formulation_residual(op, rhs, settings=st)
    out = {"workload": "Dirichlet Poisson 256^2, h = 1/256, three Gaussian bumps", "method": "FP64 Jacobi CG",
           "seconds": round(d["duration"], 5), "iterations": d["iterations"],
           "verified": bool(d["verified"]), "true_relative_residual": d["true_relative_residual"]}
    if cpu:
        base, build = cpu_lib()
        h = np.full(len(x), 1.0 / 256.0)
        base.poisson2d_cg(h, h, None, b, 1e-10, 20000, jacobi=True)  # warm-up
        t0 = time.perf_counter()
        _, it, rn, bn = base.poisson2d_cg(h, h, None, b, 1e-10, 20000, jacobi=True)
        out["cpu_baseline"] = {"seconds": round(time.perf_counter() - t0, 5), "iterations": it,
                               "relative_residual": rn / bn, "cores": base.threads(), "kind": "port",
                               "sample": f"the full solve: FP64 Jacobi CG, OpenMP, built {build}"}
    return out


def solve_magnetostatics(torch, max_iterations=20000, cpu=True):
    """192³-cell inductor (μ_r 1000 core, coil annulus): FP64 vs mixed CG to
    1e-10; FP64 Jacobi CG on the host cores beside it (a sample of
    iterations, projected to the FP64 solve's iteration count)."""
    from similie_amd import solvers as S
    from similie_amd.workloads import MU0, magnetostatics_inductor

    coords, cell_mu, cell_j = magnetostatics_inductor()
    shape = tuple(len(c) for c in coords)
    mu = S.MagnetostaticsOperator3D.nodal_mu(shape, torch.from_numpy(cell_mu).cuda(), MU0)
    op = S.MagnetostaticsOperator3D(*coords, mu, gauge_penalty=1e7)
    rhs = op.rhs(torch.from_numpy(np.ascontiguousarray(cell_j)).cuda())
    out = {"workload": "magnetostatics curl-curl, 192^3 uniform cells, mu_r 1000 core 48x48x96, coil annulus "
                       "half-widths 32-40 cells, 64 high, J = 1 A/m^2", "dofs": int(rhs.numel())}
    for prec in ("fp64", "mixed"):
        st = S.StrongFormulationSolverSettings(max_iterations=max_iterations, relative_tolerance=1e-10,
                                               precision=prec)
        print(f"bench:   magnetostatics {prec} ...", file=sys.stderr, flush=True)
        _, d = S.minimize_strong_formulation_residual(op, rhs, settings=st)
        out[prec] = {"seconds": round(d["duration"], 4), "iterations": d["iterations"],
                     "converged": bool(d["converged"]), "verified": bool(d["verified"]), "true_relative_residual": d["true_relative_residual"]}
    if out["mixed"]["seconds"] > 0:
        out["mixed_speedup"] = round(out["fp64"]["seconds"] / out["mixed"]["seconds"], 3)
    if cpu:
        base, build = cpu_lib()
        b = rhs.cpu().numpy()
        muh = mu.cpu().numpy()
        times = {}
        for k in (2, 22):
            t0 = time.perf_counter()
            base.ms3d_cg(coords, muh, 1e7, b, 1e-10, -k, jacobi=True)
            times[k] = time.perf_counter() - t0
        per = (times[22] - times[2]) / 20
        its = out["fp64"]["iterations"]
        out["cpu_baseline"] = {"seconds": round(per * its, 3), "seconds_per_iteration": round(per, 5),
                               "cores": base.threads(), "kind": "port",
                               "sample": f"FP64 Jacobi CG (OpenMP, built {build}) on the same 192^3 system: 20 "
                                         f"iterations timed (a 22-iteration run minus a 2-iteration run), projected "
                                         f"to the {its} iterations of the FP64 solve"}
    return out


def magnetostatics_assembled(torch, max_iterations=20000):
    """The 192³ inductor through the assembled path: device CSR assembly, the
    sparse product, SpdIsai generation, and FP64 CG to 1e-10 on the assembled
    matrix (Jacobi, SpdIsai) and matrix-free with SpdIsai (assembled only to
    generate the preconditioner)."""
    from similie_amd import solvers as S
    from similie_amd.workloads import MU0, magnetostatics_inductor

    coords, cell_mu, cell_j = magnetostatics_inductor()
    shape = tuple(len(c) for c in coords)
    mu = S.MagnetostaticsOperator3D.nodal_mu(shape, torch.from_numpy(cell_mu).cuda(), MU0)
    op = S.MagnetostaticsOperator3D(*coords, mu, gauge_penalty=1e7)
    rhs = op.rhs(torch.from_numpy(np.ascontiguousarray(cell_j)).cuda())
    out = {"workload": "magnetostatics inductor 192^3 cells, assembled (CSR) path"}
    A = op.assemble()  # first call: allocations and code loading
    del A
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    A = op.assemble()
    torch.cuda.synchronize()
    t_asm = time.perf_counter() - t0
    n, nnz = A.size(), A.nnz()
    csr_bytes = 12 * nnz + 4 * (n + 1)
    out["assembly"] = {"rows": n, "nnz": nnz, "seconds": round(t_asm, 5),
                       "written_GBps": round(csr_bytes / t_asm / 1e9, 1)}
    x = torch.randn(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        A.apply(x, y)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    e0.record()
    for _ in range(reps):
        A.apply(x, y)
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) / reps * 1e3
    spmv_bytes = csr_bytes + 16 * n  # values, columns, row pointers; x once, y once
    out["sparse_product"] = {"us": round(us, 1), "algorithmic_bytes": spmv_bytes,
                             "GBps": round(spmv_bytes / (us * 1e-6) / 1e9, 1),
                             "hbm_peak_fraction": round(spmv_bytes / (us * 1e-6) / 1e9 / HBM_PEAK_GBPS, 4)}
    for _ in range(5):
        op.apply(x, y)
    e0.record()
    for _ in range(reps):
        op.apply(x, y)
    e1.record()
    torch.cuda.synchronize()
    out["matrix_free_apply_us"] = round(e0.elapsed_time(e1) / reps * 1e3, 1)
    del x, y
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    G = S.spd_isai(A)
    torch.cuda.synchronize()
    out["spd_isai_generation"] = {"seconds": round(time.perf_counter() - t0, 5), "nnz": G.nnz()}
    del G
    for name, oper, pre in (("assembled_jacobi", A, "jacobi"), ("assembled_spd_isai", A, "spd_isai"),
                            ("matrix_free_spd_isai", op, "spd_isai"), ("matrix_free_jacobi", op, "jacobi")):
        print(f"bench:   {name} ...", file=sys.stderr, flush=True)
        st = S.StrongFormulationSolverSettings(max_iterations=max_iterations, relative_tolerance=1e-10,
                                               preconditioner=pre)
        _, d = S.minimize_strong_formulation_residual(oper, rhs, settings=st)
        out[name] = {"seconds": round(d["duration"], 4), "iterations": d["iterations"],
                     "converged": bool(d["converged"]), "verified": bool(d["verified"]), "true_relative_residual": d["true_relative_residual"]}
    return out


def magnetostatics_nonlinear(torch, max_iterations=20000):
    """192³-cell inductor with a nonlinear (EIcore B-H) core: Newton with
    line search to 1e-10, inner Jacobi CG in FP64 and in mixed precision."""
    from similie_amd import solvers as S
    from similie_amd.workloads import MU0, magnetostatics_nonlinear_inductor

    curve = json.loads((ROOT / "tests" / "golden" / "eicore_bh_curve.json").read_text())
    coords, cell_mu, flags, cell_j = magnetostatics_nonlinear_inductor()
    shape = tuple(len(c) for c in coords)
    mu = S.MagnetostaticsOperator3D.nodal_mu(shape, torch.from_numpy(cell_mu).cuda(), MU0)
    frac = S.NonlinearMagnetostaticsOperator3D.nodal_fraction(shape, torch.from_numpy(flags).cuda())
    op = S.NonlinearMagnetostaticsOperator3D(*coords, mu, frac, curve["b"], curve["h"], gauge_penalty=20.0)
    rhs = op.rhs(torch.from_numpy(np.ascontiguousarray(cell_j)).cuda())
    out = {"workload": "magnetostatics inductor 192^3 cells, unit spacing, EIcore B-H curve in the core "
                       "(examples/onelab_inductor/BH.pro), coil current 2000 per unit area, gauge penalty 20",
           "dofs": int(rhs.numel())}
    for prec in ("fp64", "mixed"):
        print(f"bench:   nonlinear {prec} ...", file=sys.stderr, flush=True)
        st = S.StrongFormulationSolverSettings(max_iterations=max_iterations, relative_tolerance=1e-10,
                                               precision=prec)
        _, d = S.minimize_strong_formulation_residual(op, rhs, settings=st)
        out[prec] = {"seconds": round(d["duration"], 4), "newton_steps": d["outer_iterations"],
                     "cg_iterations": d["iterations"], "converged": bool(d["converged"]), "verified": bool(d["verified"]),
                     "final_relative_residual": d["final_relative_residual"]}
    if out["mixed"]["seconds"] > 0:
        out["mixed_speedup"] = round(out["fp64"]["seconds"] / out["mixed"]["seconds"], 3)
    return out


def magnetostatics_2d(torch, max_iterations=20000):
    """2048²-cell inductor cross-section on the A_z operator: linear (μ_r 1000
    core) and nonlinear (EIcore core) solves to 1e-10, FP64 and mixed, the
    operator application timed on its own, the 2D cell post-processing."""
    from similie_amd import solvers as S
    from similie_amd.workloads import MU0, magnetostatics_2d_inductor

    curve = json.loads((ROOT / "tests" / "golden" / "eicore_bh_curve.json").read_text())
    coords, cell_mu, flags, jz = magnetostatics_2d_inductor()
    shape = tuple(len(c) for c in coords)
    nodes = shape[0] * shape[1]
    cmu, cfl, cj = (torch.from_numpy(a).cuda() for a in (cell_mu, flags, jz))
    mu = S.MagnetostaticsOperator2D.nodal_mu(shape, cmu, MU0)
    out = {"workload": "2D inductor cross-section, 2048^2 unit cells, core 512x1024 (mu_r 1000 / EIcore B-H), "
                       "coil slabs 256 wide at J_z = +-187.5", "dofs": nodes}
    op = S.MagnetostaticsOperator2D(*coords, mu)
    rhs = op.rhs(cj)
    x = torch.randn(nodes, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        op.apply(x, y)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(100):
        op.apply(x, y)
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) / 100 * 1e3
    out["apply_us"] = round(us, 2)
    out["apply_GBps"] = round(24 * nodes / (us * 1e-6) / 1e9, 1)
    for prec in ("fp64", "mixed"):
        st = S.StrongFormulationSolverSettings(max_iterations=max_iterations, relative_tolerance=1e-10,
                                               precision=prec)
        _, d = S.minimize_strong_formulation_residual(op, rhs, settings=st)
        out["linear_" + prec] = {"seconds": round(d["duration"], 4), "iterations": d["iterations"],
                                 "converged": bool(d["converged"]), "verified": bool(d["verified"]),
                                 "true_relative_residual": d["true_relative_residual"]}
    frac = S.NonlinearMagnetostaticsOperator2D.nodal_fraction(shape, cfl)
    mu0 = S.MagnetostaticsOperator2D.nodal_mu(shape, torch.full_like(cmu, MU0), MU0)
    nop = S.NonlinearMagnetostaticsOperator2D(*coords, mu0, frac, curve["b"], curve["h"])
    for prec in ("fp64", "mixed"):
        print(f"bench:   2D nonlinear {prec} ...", file=sys.stderr, flush=True)
        st = S.StrongFormulationSolverSettings(max_iterations=max_iterations, relative_tolerance=1e-10,
                                               precision=prec)
        xs, d = S.minimize_strong_formulation_residual(nop, rhs, settings=st)
        out["nonlinear_" + prec] = {"seconds": round(d["duration"], 4), "newton_steps": d["outer_iterations"],
                                    "cg_iterations": d["iterations"], "converged": bool(d["converged"]), "verified": bool(d["verified"]),
                                    "final_relative_residual": d["final_relative_residual"]}
    pp = nop.post_process(xs, torch.full_like(cmu, MU0), cfl)
    torch.cuda.synchronize()
    out["core_max_B"] = round(float(torch.max(torch.linalg.norm(pp.magnetic_induction, dim=1)[cfl > 0]).item()), 4)
    return out


def magnetostatics_post_process(torch, repeats=50):
    """Cell B, H, Maxwell stress and xy-slice force density of a potential
    on the 192³-cell inductor grid (its cell μ and nonlinear-core flags, the
    EIcore curve), timed with HIP events on the launching stream; the oracle
    on a 48³-cell sample of the same inductor beside it."""
    from similie_amd import solvers as S
    from similie_amd.workloads import MU0, magnetostatics_nonlinear_inductor

    curve = json.loads((ROOT / "tests" / "golden" / "eicore_bh_curve.json").read_text())
    coords, cell_mu, flags, _ = magnetostatics_nonlinear_inductor()
    shape = tuple(len(c) for c in coords)
    nodes, cells = int(np.prod(shape)), int(cell_mu.size)
    mu = S.MagnetostaticsOperator3D.nodal_mu(shape, torch.from_numpy(cell_mu).cuda(), MU0)
    frac = S.NonlinearMagnetostaticsOperator3D.nodal_fraction(shape, torch.from_numpy(flags).cuda())
    op = S.NonlinearMagnetostaticsOperator3D(*coords, mu, frac, curve["b"], curve["h"], gauge_penalty=20.0)
    gen = torch.Generator(device="cuda").manual_seed(5)
    pot = torch.randn(3 * nodes, generator=gen, dtype=torch.float64, device="cuda") * 1e-2
    cmu, cfl = torch.from_numpy(cell_mu).cuda(), torch.from_numpy(flags).cuda()
    for _ in range(3):
        op.post_process(pot, cmu, cfl)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(repeats):
        op.post_process(pot, cmu, cfl, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) / repeats * 1e3
    nbytes = 24 * nodes + (8 + 8 + 120) * cells
    out = {"workload": "cell post-processing of the 192^3-cell nonlinear inductor grid (B, H, Maxwell stress, "
                       "xy-slice force density), random potential",
           "cells": cells, "us_per_application": round(us, 2),
           "algorithmic_bytes": nbytes, "GBps": round(nbytes / (us * 1e-6) / 1e9, 1),
           "hbm_frac": round(nbytes / (us * 1e-6) / 1e9 / HBM_PEAK_GBPS, 4),
           "Mcells_per_s": round(cells / us, 1)}
    try:
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle_lib as orc

        sc, smu, sfl, _ = magnetostatics_nonlinear_inductor(cells=48)
        scells = int(smu.size)
        spot = np.random.default_rng(5).standard_normal(3 * int(np.prod([len(c) for c in sc]))) * 1e-2
        t0 = time.perf_counter()
        orc.ms_post_process(sc, spot, smu, sfl, curve["b"], curve["h"])
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"Mcells_per_s": round(scells / dt / 1e6, 4), "cores": 1, "kind": "port",
                               "sample": "oracle post-processing of the 48^3-cell inductor, one pass"}
    except Exception as exc:  # noqa: BLE001 - the oracle is optional beside the measurement
        out["cpu_baseline"] = {"error": f"{type(exc).__name__}: {exc}"}
    return out


def solve_uneven_batch(torch):
    """512 problems, sides log-uniform in [32, 512], one persistent launch."""
    from similie_amd.batched import BatchedPoisson2D
    from similie_amd.workloads import uneven_poisson_batch

    sides, kappa, rhs = uneven_poisson_batch()
    bp = BatchedPoisson2D(sides, kappa)
    shape = bp.autotune_shape(rhs, probe_iterations=20)
    bp.solve(rhs)  # warm-up
    _, res, secs = bp.solve(rhs)
    return {"workload": "512 Dirichlet Poisson problems, sides log-uniform in [32, 512], kappa = exp(GRF sigma 1, "
                        "corr 8), N(0,1) rhs", "method": "mixed CG (FP32 matvec, FP64 refinement), one hipGraph launch",
            "autotuned_shape": shape,
            "seconds": round(secs, 4), "converged": int(sum(bool(r["converged"]) for r in res)),
            "worst_true_relative_residual": max(r["true_relative_residual"] for r in res),
            "unknowns": int(rhs.numel())}


def solve_learned(torch, cpu=True):
    """64 held-out 256² problems (κ = exp(GRF σ 1.5, corr 16)): plain
    (unpreconditioned) FP64 CG, Jacobi FP64 CG, mixed CG and flexible CG with
    the trained network, to an FP64-verified 1e-10."""
    from similie_amd import learned as Lr
    from similie_amd.preconditioner_training import DEFAULT_WEIGHTS, evaluate

    if not DEFAULT_WEIGHTS.exists():
        return {"error": f"no trained weights at {DEFAULT_WEIGHTS.relative_to(ROOT)}"}
    net = Lr.UNet(64, 256, 256)
    net.set_params(torch.from_numpy(np.load(DEFAULT_WEIGHTS)))
    res = evaluate(net, 64, 256, 1.5, 16.0, 900000, log=lambda m: None)
    best = min(res["fcg_learned"]["seconds"], res["fcg_learned_mixed"]["seconds"])
    out = {"workload": "64 held-out Poisson problems 256^2, kappa = exp(GRF sigma 1.5, corr 16), N(0,1) rhs",
           "methods": res,
           "learned_speedup_vs_plain_fp64_cg": round(res["cg_plain"]["seconds"] / best, 3),
           "learned_speedup_vs_jacobi_fp64_cg": round(res["cg"]["seconds"] / best, 3)}
    if cpu:
        from similie_amd.preconditioner_training import make_problems

        base, build = cpu_lib()
        kappa, b = make_problems(64, 256, 1.5, 16.0, 900000)
        kh, bh = kappa[:4].cpu().numpy(), b[:4].cpu().numpy()
        ones = np.ones(256)
        t0 = time.perf_counter()
        its = [base.poisson2d_cg(ones, ones, kh[p], bh[p], 1e-10, 20000, jacobi=True)[1] for p in range(4)]
        secs = time.perf_counter() - t0
        out["cpu_baseline"] = {"seconds": round(secs * 16, 3), "iterations_mean": float(np.mean(its)),
                               "cores": base.threads(), "kind": "port",
                               "sample": f"FP64 Jacobi CG (OpenMP, built {build}) on the first 4 of the 64 problems "
                                         f"to 1e-10, their time x 16"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=6.0,
                    help="time budget of each CPU Laplacian baseline (at least 5 applications)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-solves", action="store_true", help="skip the solver workloads")
    ap.add_argument("--solves", default="", help="comma-separated subset of the solver workloads to run")
    ap.add_argument("--ms-iterations", type=int, default=20000,
                    help="iteration cap of each magnetostatics solve")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this build has one MI355X, and the Laplacian "
              "is not sharded across GPUs (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch

    from similie_amd import exterior as E

    torch.cuda.set_device(0)
    grid, c, g, w, out = workload(torch, E)
    staged = E.make_staged_laplacian(grid, 1, out, w)
    # the kernel shape is picked by the library's autotuner on these arrays,
    # before the timed region
    tile = staged.autotune(out, w, repeats=5)
    for _ in range(args.warmup):
        staged.run(out, w)
    torch.cuda.synchronize()

    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record()
    for _ in range(args.steps):
        staged.run(out, w)
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = start.elapsed_time(stop) / args.steps

    B = algorithmic_bytes()
    value = B * args.steps / wall / 1e9
    achieved = B / (kernel_ms * 1e-3) / 1e9

    # plain copy of the same byte count for reference (torch's copy kernel)
    a = torch.ones(B // 16, dtype=torch.float64, device="cuda")
    b2 = torch.empty_like(a)
    for _ in range(3):
        b2.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        b2.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    copy_gbps = 2 * a.numel() * 8 / (e0.elapsed_time(e1) / 20 * 1e-3) / 1e9
    del a, b2
    # the Laplacian's byte mix as plain streaming: read two arrays, write one
    # (torch's elementwise add), B bytes in total
    a1 = torch.ones(B // 24, dtype=torch.float64, device="cuda")
    a2 = torch.ones_like(a1)
    c3 = torch.empty_like(a1)
    for _ in range(3):
        torch.add(a1, a2, out=c3)
    e0.record()
    for _ in range(20):
        torch.add(a1, a2, out=c3)
    e1.record()
    torch.cuda.synchronize()
    mix_gbps = 3 * a1.numel() * 8 / (e0.elapsed_time(e1) / 20 * 1e-3) / 1e9
    del a1, a2, c3

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "3D Hodge Laplacian (δd+dδ) of a 1-form, 256^3 nodes, geometric spacing ratio 1.01 "
                        "on [0,1]^3, diagonal metric g_dd = 1 + x_d^2, N(0,1) edge values, output on 255^3 nodes",
            "grid": [N, N, N],
            "degree": 1,
            "algorithmic_bytes_per_step": B,
            "hbm_peak_fraction": round(value / HBM_PEAK_GBPS, 4),
            "torch_copy_GBps": round(copy_gbps, 1),
            "torch_read2_write1_GBps": round(mix_gbps, 1),
            "kernel_shape": tile,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": profiled_traffic(tile),
        },
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline(args.cpu_seconds),
    }
    if not args.no_solves:
        del grid, g, w, out, staged
        torch.cuda.empty_cache()
        cpu = not args.no_cpu_baseline
        solves = {
            "laplacian_2d_0form": lambda: laplacian_2d_0form(torch),
            "poisson2d_bumps": lambda: solve_poisson_bumps(torch, cpu),
            "uneven_batch_512": lambda: solve_uneven_batch(torch),
            "learned_preconditioner": lambda: solve_learned(torch, cpu),
            "magnetostatics_inductor": lambda: solve_magnetostatics(torch, args.ms_iterations, cpu),
            "magnetostatics_assembled": lambda: magnetostatics_assembled(torch, args.ms_iterations),
            "magnetostatics_nonlinear": lambda: magnetostatics_nonlinear(torch, args.ms_iterations),
            "magnetostatics_post_process": lambda: magnetostatics_post_process(torch),
            "magnetostatics_2d": lambda: magnetostatics_2d(torch, args.ms_iterations),
        }
        pick = [k for k in args.solves.split(",") if k]
        unknown = [k for k in pick if k not in solves]
        if unknown:
            raise SystemExit(f"unknown solver workloads {unknown}; known: {sorted(solves)}")
        line["solves"] = {k: _guard(k, f) for k, f in solves.items() if not pick or k in pick}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
