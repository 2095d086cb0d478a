"""Batched persistent Poisson solves (many uneven 2D problems, one launch).

No reference counterpart: the reference solves one operator per Ginkgo
call.  See sil_batched_poisson2d_* in include/sil_hip.h.
"""
import ctypes

import torch

from . import _lib as L
from .solvers import StrongFormulationSolverSettings


class BatchedPoisson2D:
    """Dirichlet-zero d0ᵀ ⋆1^κ d0 problems on unit-spaced n×n grids.

    sides: problem sizes (nodes per axis); kappa: concatenated nodal
    coefficients (float64 CUDA tensor), problem p occupying
    [offset(p), offset(p) + sides[p]²) with x slowest."""

    def __init__(self, sides, kappa):
        self.sides = [int(s) for s in sides]
        if kappa.dtype != torch.float64 or not kappa.is_cuda or not kappa.is_contiguous():
            raise TypeError("kappa must be a contiguous float64 CUDA tensor")
        self.total = sum(s * s for s in self.sides)
        if kappa.numel() != self.total:
            raise ValueError(f"kappa has {kappa.numel()} entries, the problems need {self.total}")
        arr = (ctypes.c_int32 * len(self.sides))(*self.sides)
        h = ctypes.c_void_p()
        L.check(L.lib().sil_batched_poisson2d_create(ctypes.byref(h), len(self.sides), arr, L.ptr(kappa)))
        self.handle = h
        self._destroy = L.lib().sil_batched_poisson2d_destroy
        self.block = 256

    def offset(self, p):
        return int(L.lib().sil_batched_poisson2d_offset(self.handle, p))

    def set_block(self, block):
        L.check(L.lib().sil_batched_poisson2d_set_block(self.handle, int(block)))
        self.block = int(block)

    def autotune(self, rhs, probe_iterations=20):
        """Times every workgroup size on a fixed number of iterations of every
        problem and keeps the fastest; returns (block, seconds per candidate)."""
        chosen = ctypes.c_int32()
        secs = (ctypes.c_double * 3)()
        L.check(L.lib().sil_batched_poisson2d_autotune(self.handle, L.ptr(rhs), int(probe_iterations),
                                                       ctypes.byref(chosen), secs))
        self.block = chosen.value
        return chosen.value, {256: secs[0], 512: secs[1], 1024: secs[2]}

    def autotune_shape(self, rhs, probe_iterations=20):
        """the wider autotuner (sil_batched_poisson2d_autotune_shape): the
        wide-path threshold, the workgroups per wide problem and the
        persistent workgroup size; returns the chosen shape"""
        sh = L.BatchShape()
        L.check(L.lib().sil_batched_poisson2d_autotune_shape(self.handle, L.ptr(rhs), int(probe_iterations),
                                                             ctypes.byref(sh)))
        self.block = sh.block
        return {"wide_side": sh.wide_side, "wide_workgroups": sh.wide_workgroups, "block": sh.block,
                "wide_problems": sh.wide_problems, "probe_seconds": sh.seconds}

    def solve(self, rhs, solution=None, settings=None, stream=None):
        settings = settings or StrongFormulationSolverSettings(relative_tolerance=1e-10, precision="mixed",
                                                               preconditioner="jacobi", max_iterations=20000)
        solution = torch.empty_like(rhs) if solution is None else solution
        res = (L.BatchResult * len(self.sides))()
        secs = ctypes.c_double()
        cs = settings.to_c()
        L.check(L.lib().sil_batched_poisson2d_solve(self.handle, L.ptr(rhs), L.ptr(solution), ctypes.byref(cs), res,
                                                    ctypes.byref(secs), L.stream_handle(stream)))
        results = [{"iterations": r.iterations, "converged": r.converged,
                    "true_relative_residual": r.true_relative_residual} for r in res]
        return solution, results, secs.value

    def __del__(self):
        # the destroy function was bound at construction: module globals may
        # already be gone when this runs at interpreter exit
        h = getattr(self, "handle", None)
        fn = getattr(self, "_destroy", None)
        if h is not None and h.value and fn is not None:
            fn(h)
            self.handle = None
