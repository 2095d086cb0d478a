"""sil::exterior on MI355X: Python mirror of the reference's operator calls.

Each function keeps the argument meaning of the reference entry point it
names (src/similie/exterior/*.hpp) and forwards to libsil_hip.so on the
caller's torch stream.  Forms are torch float64 CUDA tensors; by default a
k-form on an N-dimensional grid is index-last, shape (n0, ..., n_{N-1}, C(N,k))
— the reference's DiscreteDomain<DDim..., Index> with layout_right — and the
component order is the lexicographic order of the k-subsets of the axes.
`grid_dims`, `comp_dim` and `batch_dims` describe other layouts (index-first,
spectator dimensions).
"""
import ctypes
from math import comb

import torch

from . import _lib as L


def make_view(t, grid_dims=None, comp_dim=-1, batch_dims=()):
    """sil_view of tensor `t`: `grid_dims` are the form axes in order, `comp_dim`
    the component dimension (None for a missing component axis), `batch_dims`
    the spectator dimensions."""
    nd = t.dim()
    comp = None if comp_dim is None else comp_dim % nd
    batch = [d % nd for d in batch_dims]
    if grid_dims is None:
        grid_dims = [d for d in range(nd) if d != comp and d not in batch]
    grid = [d % nd for d in grid_dims]
    v = L.View()
    v.ndim = len(grid)
    v.nbatch = len(batch)
    for a, d in enumerate(grid):
        v.extent[a] = t.shape[d]
        v.stride[a] = t.stride(d)
    v.comp_stride = t.stride(comp) if comp is not None else 0
    v.comp_extent = t.shape[comp] if comp is not None else 1
    for a, d in enumerate(batch):
        v.batch_extent[a] = t.shape[d]
        v.batch_stride[a] = t.stride(d)
    return v


def _check_tensor(t, name):
    if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or not t.is_cuda:
        raise TypeError(f"{name} must be a float64 CUDA tensor")


class Grid:
    """Node grid geometry: the reference's `position` and `metric` tensors.

    axis_coords: per-axis 1-D coordinate sequences (rectilinear grid), or
    nodal_positions: CUDA tensor (n0, ..., n_{N-1}, N).
    metric: None (identity), CUDA tensor (..., N) for a diagonal metric or
    (..., N(N+1)/2) for a symmetric one (order XX XY XZ YY YZ ZZ).
    """

    def __init__(self, extents, axis_coords=None, nodal_positions=None, metric=None, metric_kind=None):
        self.extents = [int(e) for e in extents]
        self.ndim = len(self.extents)
        self._keep = []
        if metric is None:
            kind = L.METRIC_IDENTITY
        else:
            _check_tensor(metric, "metric")
            metric = metric.contiguous()
            ncomp = metric.shape[-1]
            if metric_kind is not None:
                kind = metric_kind
            elif ncomp == self.ndim and self.ndim > 1:
                kind = L.METRIC_DIAGONAL
            else:
                kind = L.METRIC_SYMMETRIC if ncomp == self.ndim * (self.ndim + 1) // 2 else L.METRIC_DIAGONAL
        self.metric = metric
        self.metric_kind = kind
        ext = (ctypes.c_int64 * self.ndim)(*self.extents)
        handle = ctypes.c_void_p()
        if axis_coords is not None:
            arrs = [
                (ctypes.c_double * len(c))(*[float(v) for v in (c.tolist() if hasattr(c, "tolist") else c)])
                for c in axis_coords
            ]
            self._keep = arrs
            ptrs = (ctypes.c_void_p * self.ndim)(*[ctypes.cast(a, ctypes.c_void_p) for a in arrs])
            L.check(L.lib().sil_grid_create(ctypes.byref(handle), self.ndim, ext, ptrs, None, kind, L.ptr(metric)))
        else:
            _check_tensor(nodal_positions, "nodal_positions")
            self.nodal = nodal_positions.contiguous()
            L.check(L.lib().sil_grid_create(ctypes.byref(handle), self.ndim, ext, None, L.ptr(self.nodal), kind,
                                            L.ptr(metric)))
        self.handle = handle
        self._destroy = L.lib().sil_grid_destroy

    def __del__(self):
        # the destroy function was bound at construction: module globals may
        # already be gone when this runs at interpreter exit
        h = getattr(self, "handle", None)
        fn = getattr(self, "_destroy", None)
        if h is not None and h.value and fn is not None:
            fn(h)
            self.handle = None


def _views(out, tensor, grid_dims, comp_dim, batch_dims):
    _check_tensor(out, "out")
    _check_tensor(tensor, "tensor")
    return (make_view(tensor, grid_dims, comp_dim, batch_dims), make_view(out, grid_dims, comp_dim, batch_dims))


def coboundary(out, tensor, degree, *, grid_dims=None, comp_dim=-1, batch_dims=(), stream=None):
    """deriv / coboundary (coboundary.hpp:614-654, 698-709): `tensor` is a
    degree-k form, `out` the (k+1)-form; samples beyond the input domain are
    clamped."""
    vin, vout = _views(out, tensor, grid_dims, comp_dim, batch_dims)
    L.check(L.lib().sil_coboundary(degree, L.ptr(tensor), ctypes.byref(vin), L.ptr(out), ctypes.byref(vout),
                                   L.stream_handle(stream)))
    return out


deriv = coboundary


def transposed_coboundary(out, tensor, degree, *, grid_dims=None, comp_dim=-1, batch_dims=(), stream=None):
    """transposed_coboundary (coboundary.hpp:656-696): backward differences,
    zero outside the input domain, negated."""
    vin, vout = _views(out, tensor, grid_dims, comp_dim, batch_dims)
    L.check(L.lib().sil_transposed_coboundary(degree, L.ptr(tensor), ctypes.byref(vin), L.ptr(out),
                                              ctypes.byref(vout), L.stream_handle(stream)))
    return out


def fill_discrete_hodge_star(grid, degree, hodge_star=None, stream=None):
    """fill_discrete_hodge_star<I(degree indices), J(N-degree indices)>
    (hodge_star.hpp:453-494): per-node coefficients, shape
    (*extents, N**degree, C(N, N-degree))."""
    n = grid.ndim
    shape = tuple(grid.extents) + (n ** degree, comb(n, n - degree))
    if hodge_star is None:
        hodge_star = torch.empty(shape, dtype=torch.float64, device="cuda")
    _check_tensor(hodge_star, "hodge_star")
    if tuple(hodge_star.shape) != shape or not hodge_star.is_contiguous():
        raise ValueError(f"hodge_star must be a contiguous tensor of shape {shape}")
    L.check(L.lib().sil_fill_discrete_hodge_star(grid.handle, degree, L.ptr(hodge_star), L.stream_handle(stream)))
    return hodge_star


def hodge_star(grid, degree, out, tensor, *, comp_dim=-1, stream=None):
    """Discrete ⋆ of a degree-k form onto the dual (N-k)-form (the tensor_prod
    of the filled star with the form, codifferential.hpp:584)."""
    vin, vout = _views(out, tensor, None, comp_dim, ())
    L.check(L.lib().sil_hodge_star(grid.handle, degree, L.ptr(tensor), ctypes.byref(vin), L.ptr(out),
                                   ctypes.byref(vout), L.stream_handle(stream)))
    return out


def codifferential(grid, degree, out, tensor, *, comp_dim=-1, stream=None):
    """codifferential (codifferential.hpp:551-656): degree-k form on the grid's
    node domain → (k-1)-form on out's extents."""
    vin, vout = _views(out, tensor, None, comp_dim, ())
    L.check(L.lib().sil_codifferential(grid.handle, degree, L.ptr(tensor), ctypes.byref(vin), L.ptr(out),
                                       ctypes.byref(vout), L.stream_handle(stream)))
    return out


def laplacian(grid, degree, out, tensor, *, comp_dim=-1, stream=None):
    """laplacian (laplacian.hpp:814-951): δd + dδ of a degree-k form on the
    grid's node domain, written on out's extents (the reference's callers use
    the node domain minus the last node of every axis)."""
    vin, vout = _views(out, tensor, None, comp_dim, ())
    L.check(L.lib().sil_laplacian(grid.handle, degree, L.ptr(tensor), ctypes.byref(vin), L.ptr(out),
                                  ctypes.byref(vout), L.stream_handle(stream)))
    return out


class StagedLaplacian:
    """make_staged_laplacian(exec, out, in, metric, position).run(out, in)
    (laplacian.hpp:312-358)."""

    def __init__(self, grid, degree, out, tensor, comp_dim=-1):
        vin, vout = _views(out, tensor, None, comp_dim, ())
        self.grid = grid
        # the handle keeps the strides and the kernel path chosen here
        self.layout_in = (tuple(tensor.shape), tuple(tensor.stride()))
        self.layout_out = (tuple(out.shape), tuple(out.stride()))
        h = ctypes.c_void_p()
        L.check(L.lib().sil_make_staged_laplacian(ctypes.byref(h), grid.handle, degree, ctypes.byref(vin),
                                                  ctypes.byref(vout)))
        self.handle = h
        self._destroy = L.lib().sil_staged_laplacian_destroy

    def run(self, out, tensor, stream=None):
        _check_tensor(out, "out")
        _check_tensor(tensor, "tensor")
        if (tuple(out.shape), tuple(out.stride())) != self.layout_out or \
                (tuple(tensor.shape), tuple(tensor.stride())) != self.layout_in:
            raise ValueError("staged laplacian called with shapes or strides it was not built for")
        L.check(L.lib().sil_staged_laplacian_run(self.handle, L.ptr(out), L.ptr(tensor), L.stream_handle(stream)))
        return out

    def autotune(self, out, tensor, repeats=5, stream=None):
        """times the fused 3D kernel shapes on these arrays and keeps the
        fastest (sil_staged_laplacian_autotune); returns the chosen shape"""
        _check_tensor(out, "out")
        _check_tensor(tensor, "tensor")
        if (tuple(out.shape), tuple(out.stride())) != self.layout_out or \
                (tuple(tensor.shape), tuple(tensor.stride())) != self.layout_in:
            raise ValueError("staged laplacian called with shapes or strides it was not built for")
        t = L.Lap3dTile()
        L.check(L.lib().sil_staged_laplacian_autotune(self.handle, L.ptr(out), L.ptr(tensor), int(repeats),
                                                      ctypes.byref(t), L.stream_handle(stream)))
        names = {0: "direct", 1: "march", 2: "pipe", 3: "stage"}
        return {"variant": names.get(t.variant, str(t.variant)), "rows": t.rows, "xchunk": t.xchunk,
                "lanes_or_depth": t.lanes_or_depth, "nontemporal": bool(t.nontemporal),
                "microseconds": t.microseconds}

    def __del__(self):
        # the destroy function was bound at construction: module globals may
        # already be gone when this runs at interpreter exit
        h = getattr(self, "handle", None)
        fn = getattr(self, "_destroy", None)
        if h is not None and h.value and fn is not None:
            fn(h)
            self.handle = None


def make_staged_laplacian(grid, degree, out, tensor, comp_dim=-1):
    return StagedLaplacian(grid, degree, out, tensor, comp_dim)


class StagedCodifferential:
    """StagedCodifferential(exec, tensor, metric, position).run(out, tensor)
    (codifferential.hpp:321-508): a degree-k form → its codifferential on
    out's extents, validated once and run on arrays of the same layout."""

    def __init__(self, grid, degree, out, tensor, comp_dim=-1):
        vin, vout = _views(out, tensor, None, comp_dim, ())
        self.grid = grid
        self.layout_in = (tuple(tensor.shape), tuple(tensor.stride()))
        self.layout_out = (tuple(out.shape), tuple(out.stride()))
        h = ctypes.c_void_p()
        L.check(L.lib().sil_make_staged_codifferential(ctypes.byref(h), grid.handle, degree, ctypes.byref(vin),
                                                       ctypes.byref(vout)))
        self.handle = h
        self._destroy = L.lib().sil_staged_codifferential_destroy

    def run(self, out, tensor, stream=None):
        _check_tensor(out, "out")
        _check_tensor(tensor, "tensor")
        if (tuple(out.shape), tuple(out.stride())) != self.layout_out or \
                (tuple(tensor.shape), tuple(tensor.stride())) != self.layout_in:
            raise ValueError("staged codifferential called with shapes or strides it was not built for")
        L.check(L.lib().sil_staged_codifferential_run(self.handle, L.ptr(out), L.ptr(tensor),
                                                      L.stream_handle(stream)))
        return out

    def __del__(self):
        h = getattr(self, "handle", None)
        fn = getattr(self, "_destroy", None)
        if h is not None and h.value and fn is not None:
            fn(h)
            self.handle = None


def make_staged_codifferential(grid, degree, out, tensor, comp_dim=-1):
    return StagedCodifferential(grid, degree, out, tensor, comp_dim)
