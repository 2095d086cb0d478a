"""Linear operators and the strong-formulation CG solve on MI355X.

Python mirror of the reference's solver boundary
(src/similie/solvers/minimize_strong_formulation_residual.hpp:119-144,
1308-1382) and of the 3D magnetostatics operator
(onelab_interface/magnetostatics_onelab.hpp:1524-2789).  Vectors are torch
float64 CUDA tensors; everything runs through libsil_hip.so.
"""
import ctypes
from dataclasses import dataclass
from typing import NamedTuple

import numpy as np
import torch

from . import _lib as L


def _f64(t, name):
    if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
        raise TypeError(f"{name} must be a contiguous float64 CUDA tensor")
    return t


def _host_array(seq):
    vals = seq.tolist() if hasattr(seq, "tolist") else list(seq)
    return (ctypes.c_double * len(vals))(*[float(v) for v in vals])


class LinearOperator:
    """A sil_operator handle: apply (FP64), apply_f32, diagonal."""

    handle = None

    def size(self):
        return int(L.lib().sil_operator_size(self.handle))

    def apply(self, x, out=None, stream=None):
        _f64(x, "x")
        out = torch.empty_like(x) if out is None else _f64(out, "out")
        L.check(L.lib().sil_operator_apply(self.handle, L.ptr(x), L.ptr(out), L.stream_handle(stream)))
        return out

    def apply_f32(self, x, out=None, stream=None):
        if x.dtype != torch.float32 or not x.is_cuda:
            raise TypeError("x must be a float32 CUDA tensor")
        out = torch.empty_like(x) if out is None else out
        L.check(L.lib().sil_operator_apply_f32(self.handle, L.ptr(x), L.ptr(out), L.stream_handle(stream)))
        return out

    def diagonal(self, stream=None):
        d = torch.empty(self.size(), dtype=torch.float64, device="cuda")
        L.check(L.lib().sil_operator_diagonal(self.handle, L.ptr(d), L.stream_handle(stream)))
        return d

    def __del__(self):
        # the destroy function was bound at construction: module globals may
        # already be gone when this runs at interpreter exit
        h = getattr(self, "handle", None)
        fn = getattr(self, "_destroy", None)
        if h is not None and h.value and fn is not None:
            fn(h)
            self.handle = None


class CellPostProcessFields(NamedTuple):
    """per-cell post-processing outputs (CellPostProcessFields,
    magnetostatics_onelab.hpp:402-408)"""

    magnetic_induction: torch.Tensor
    magnetic_field: torch.Tensor
    maxwell_stress: torch.Tensor
    force_density: torch.Tensor


class MagnetostaticsOperator3D(LinearOperator):
    """MagnetostaticsOperator3D with precomputed-stencil semantics: DOF order
    3·(i + nx·(j + ny·k)) + c, identity rows on boundary nodes, gauge
    penalty P·w^A on interior rows."""

    def __init__(self, x_coords, y_coords, z_coords, mu_nodal, gauge_penalty=20.0):
        self.shape = (len(x_coords), len(y_coords), len(z_coords))
        _f64(mu_nodal, "mu_nodal")
        h = ctypes.c_void_p()
        xs, ys, zs = _host_array(x_coords), _host_array(y_coords), _host_array(z_coords)
        nx, ny, nz = self.shape
        L.check(L.lib().sil_magnetostatics3d_create(ctypes.byref(h), nx, ny, nz, ctypes.cast(xs, ctypes.c_void_p),
                                                    ctypes.cast(ys, ctypes.c_void_p), ctypes.cast(zs, ctypes.c_void_p),
                                                    L.ptr(mu_nodal), float(gauge_penalty)))
        self.handle = h
        self._destroy = L.lib().sil_operator_destroy

    @staticmethod
    def nodal_mu(shape, cell_mu, mu0, stream=None):
        """mean of the adjacent cells' μ (magnetostatics_onelab.hpp:5029-5066)"""
        nx, ny, nz = shape
        _f64(cell_mu, "cell_mu")
        mu = torch.empty(nx * ny * nz, dtype=torch.float64, device="cuda")
        L.check(L.lib().sil_magnetostatics3d_nodal_mu(nx, ny, nz, L.ptr(cell_mu), float(mu0), L.ptr(mu),
                                                      L.stream_handle(stream)))
        return mu

    def rhs(self, cell_current, stream=None):
        """⅛·|dual face|·J over the cells around each interior node
        (magnetostatics_onelab.hpp:4919-4975); cell_current: (cells, 3)."""
        _f64(cell_current, "cell_current")
        b = torch.empty(self.size(), dtype=torch.float64, device="cuda")
        L.check(L.lib().sil_magnetostatics3d_rhs(self.handle, L.ptr(cell_current), L.ptr(b), L.stream_handle(stream)))
        return b

    def post_process(self, potential, cell_mu, cell_nonlinear=None, stream=None):
        """CellPostProcessFields of every cell (magnetostatics_onelab.hpp
        :3984-4046, 4048-4168, 4215-4254): magnetic_induction (cells, 3),
        magnetic_field (cells, 3), maxwell_stress (cells, 6: xx yy zz xy xz
        yz) and force_density (cells, 3), cells x fastest.  cell_nonlinear
        (0/1 per cell) selects the B-H curve and needs a nonlinear operator."""
        _f64(potential, "potential")
        _f64(cell_mu, "cell_mu")
        if cell_nonlinear is not None:
            _f64(cell_nonlinear, "cell_nonlinear")
        nx, ny, nz = self.shape
        cells = max(nx - 1, 0) * max(ny - 1, 0) * max(nz - 1, 0)
        out = [torch.empty((cells, w), dtype=torch.float64, device="cuda") for w in (3, 3, 6, 3)]
        L.check(L.lib().sil_magnetostatics3d_post_process(
            self.handle, L.ptr(potential), L.ptr(cell_mu), L.ptr(cell_nonlinear) if cell_nonlinear is not None else None,
            *[L.ptr(o) for o in out], L.stream_handle(stream)))
        return CellPostProcessFields(*out)


    def set_switches(self, y_response_sign=1.0, divergence_gauge=False, gauge_sign=1.0, criterion="moments"):
        """the operator's switches (magnetostatics_onelab.hpp:796-828) and
        the right-hand side's criterion (:4930-4935): y_response_sign −1 flips
        the y moment's response; divergence_gauge replaces the penalty term
        P·w^A·A by (P · gauge_sign)·δᵀWδA; criterion "potential" gives a zero
        right-hand side."""
        sw = L.Ms3dSwitches(float(y_response_sign), int(bool(divergence_gauge)), float(gauge_sign),
                            CRITERIA[criterion])
        L.check(L.lib().sil_magnetostatics3d_set_switches(self.handle, ctypes.byref(sw)))
        return self

    def assemble(self, stream=None):
        """assemble_matrix_data + csr_from_matrix_data
        (magnetostatics_onelab.hpp:3115-3536,
        minimize_strong_formulation_residual.hpp:565-646): the same operator
        as a CSR matrix assembled on the device."""
        h = ctypes.c_void_p()
        L.check(L.lib().sil_magnetostatics3d_assemble(self.handle, ctypes.byref(h), L.stream_handle(stream)))
        return AssembledOperator._adopt(h)


class NonlinearMagnetostaticsOperator3D(MagnetostaticsOperator3D):
    """MagnetostaticsOperator3D with MaterialMagnetostaticsHamiltonian
    (magnetostatics_onelab.hpp:957-1011): nodes whose ferromagnetic fraction
    exceeds 0.5 follow the interpolated B-H curve (bh_b, bh_h; ≤ 64 samples,
    nonlinear_magnetostatics.py:14-227), the others are linear with mu_nodal.
    apply(x) is F(x); minimize_strong_formulation_residual runs Newton
    (minimize_strong_formulation_residual.hpp:1383-1569)."""

    def __init__(self, x_coords, y_coords, z_coords, mu_nodal, ferromagnetic_nodal, bh_b, bh_h,
                 gauge_penalty=20.0):
        self.shape = (len(x_coords), len(y_coords), len(z_coords))
        _f64(mu_nodal, "mu_nodal")
        _f64(ferromagnetic_nodal, "ferromagnetic_nodal")
        if len(bh_b) != len(bh_h):
            raise ValueError("bh_b and bh_h must have the same length")
        h = ctypes.c_void_p()
        xs, ys, zs = _host_array(x_coords), _host_array(y_coords), _host_array(z_coords)
        bb, hh = _host_array(bh_b), _host_array(bh_h)
        nx, ny, nz = self.shape
        L.check(L.lib().sil_magnetostatics3d_nonlinear_create(
            ctypes.byref(h), nx, ny, nz, ctypes.cast(xs, ctypes.c_void_p), ctypes.cast(ys, ctypes.c_void_p),
            ctypes.cast(zs, ctypes.c_void_p), L.ptr(mu_nodal), L.ptr(ferromagnetic_nodal),
            ctypes.cast(bb, ctypes.c_void_p), ctypes.cast(hh, ctypes.c_void_p), len(bh_b), float(gauge_penalty)))
        self.handle = h
        self._destroy = L.lib().sil_operator_destroy

    @staticmethod
    def nodal_fraction(shape, cell_flags, stream=None):
        """mean over the adjacent cells of a 0/1 nonlinear-material flag
        (magnetostatics_onelab.hpp:5026-5063; 0 where no cell is adjacent)"""
        return MagnetostaticsOperator3D.nodal_mu(shape, cell_flags, 0.0, stream)

    def apply_jacobian(self, state, v, out=None, stream=None):
        """apply_jacobian (magnetostatics_onelab.hpp:2918-3113): J(state)·v"""
        _f64(state, "state")
        _f64(v, "v")
        out = torch.empty_like(v) if out is None else _f64(out, "out")
        L.check(L.lib().sil_magnetostatics3d_apply_jacobian(self.handle, L.ptr(state), L.ptr(v), L.ptr(out),
                                                             L.stream_handle(stream)))
        return out

    def assemble_jacobian(self, state, stream=None):
        """assemble_matrix_data(model, state) (magnetostatics_onelab.hpp
        :3115-3536): the Jacobian at `state` as a CSR matrix assembled on the
        device."""
        _f64(state, "state")
        if state.numel() != self.size():
            raise ValueError(f"state must hold {self.size()} values")
        h = ctypes.c_void_p()
        L.check(L.lib().sil_magnetostatics3d_assemble_jacobian(self.handle, L.ptr(state), ctypes.byref(h),
                                                               L.stream_handle(stream)))
        return AssembledOperator._adopt(h)

    def jacobian_diagonal(self, state, stream=None):
        _f64(state, "state")
        d = torch.empty(self.size(), dtype=torch.float64, device="cuda")
        L.check(L.lib().sil_magnetostatics3d_jacobian_diagonal(self.handle, L.ptr(state), L.ptr(d),
                                                               L.stream_handle(stream)))
        return d


class AssembledOperator(LinearOperator):
    """gko::matrix::Csr<double, int32>: a square CSR matrix held on the
    device (int32 row pointers and columns, ascending within a row, FP64
    values); apply is the sparse product."""

    @classmethod
    def _adopt(cls, handle):
        op = cls.__new__(cls)
        op.handle = handle
        op._destroy = L.lib().sil_operator_destroy
        return op

    @classmethod
    def from_csr(cls, row_ptr, cols, vals, stream=None):
        """Csr::create(exec, size, values, columns, row_ptrs)
        (minimize_strong_formulation_residual.hpp:613-620) from int32 / int32
        / float64 tensors (host or CUDA; copied)."""
        rp = row_ptr.to(torch.int32).contiguous()
        cc = cols.to(torch.int32).contiguous()
        vv = vals.to(torch.float64).contiguous()
        h = ctypes.c_void_p()
        L.check(L.lib().sil_csr_create(ctypes.byref(h), rp.numel() - 1, cc.numel(), ctypes.c_void_p(rp.data_ptr()),
                                       ctypes.c_void_p(cc.data_ptr()), ctypes.c_void_p(vv.data_ptr()),
                                       L.stream_handle(stream)))
        return cls._adopt(h)

    def nnz(self):
        return int(L.lib().sil_csr_nnz(self.handle))

    def csr(self, stream=None):
        """(row_ptr int32, cols int32, vals float64) CUDA tensors"""
        n, nnz = self.size(), self.nnz()
        rp = torch.empty(n + 1, dtype=torch.int32, device="cuda")
        cc = torch.empty(max(nnz, 1), dtype=torch.int32, device="cuda")
        vv = torch.empty(max(nnz, 1), dtype=torch.float64, device="cuda")
        L.check(L.lib().sil_csr_copy(self.handle, ctypes.c_void_p(rp.data_ptr()), ctypes.c_void_p(cc.data_ptr()),
                                     ctypes.c_void_p(vv.data_ptr()), L.stream_handle(stream)))
        return rp, cc[:nnz], vv[:nnz]


def spd_isai(operator, stream=None):
    """gko::preconditioner::SpdIsai::generate (minimize_strong_formulation_
    residual.hpp:963-967): the factor G of M = Gᵀ G ≈ A⁻¹ (pattern of tril(A))
    as an AssembledOperator; `operator` is assembled or the magnetostatics
    operator."""
    h = ctypes.c_void_p()
    L.check(L.lib().sil_spd_isai_create(operator.handle, ctypes.byref(h), L.stream_handle(stream)))
    return AssembledOperator._adopt(h)


CRITERIA = {"potential": 0, "moments": 1, "potential_and_moments": 2}


def diagnostic_cell_mask(cell_tags, diagnostic_region_tags=(), positive_conductor_tags=()):
    """the per-cell mask of the air-gap diagnostics from physical tags: bit 0
    in a diagnostic region, bit 1 a positively oriented conductor (device
    uint8)"""
    tags = np.asarray(cell_tags if not torch.is_tensor(cell_tags) else cell_tags.cpu().numpy())
    mask = np.isin(tags, list(diagnostic_region_tags)).astype(np.uint8)
    mask |= (np.isin(tags, list(positive_conductor_tags)).astype(np.uint8) << 1)
    return torch.from_numpy(mask).cuda()


def air_gap_diagnostics(operator, fields, cell_mask, cell_current, num_turns=1.0, length_z=None, z_value=0.0,
                        stream=None):
    """the reference's air-gap diagnostics (magnetostatics_onelab.hpp
    :4666-4720 2D, 5246-5307 3D, 328-388 for the published quantities):
    upper air-gap faces of the diagnostic region, flux Φ, Maxwell traction
    integral (x-mirror projected when the faces are symmetric) and mean
    magnitude, the positive conductor's current and the inductance
    Φ / (N I).  fields: CellPostProcessFields of the solution; cell_mask:
    diagnostic_cell_mask; cell_current: (cells, 3) in 3D, (cells,) J_z in 2D;
    length_z and z_value: the 2D extrusion length and face z."""
    if cell_mask.dtype != torch.uint8 or not cell_mask.is_cuda or not cell_mask.is_contiguous():
        raise ValueError("cell_mask must be a contiguous uint8 CUDA tensor")
    b, h = _f64(fields.magnetic_induction, "magnetic_induction"), _f64(fields.magnetic_field, "magnetic_field")
    jc = _f64(cell_current, "cell_current")
    r = L.AirGapResult()
    if isinstance(operator, MagnetostaticsOperator2D):
        if length_z is None:
            raise ValueError("a 2D diagnostic needs length_z")
        L.check(L.lib().sil_magnetostatics2d_air_gap_diagnostics(
            operator.handle, L.ptr(b), L.ptr(h), L.ptr(cell_mask), L.ptr(jc), float(length_z), float(z_value),
            ctypes.byref(r), L.stream_handle(stream)))
    else:
        L.check(L.lib().sil_magnetostatics3d_air_gap_diagnostics(
            operator.handle, L.ptr(b), L.ptr(h), L.ptr(cell_mask), L.ptr(jc), ctypes.byref(r),
            L.stream_handle(stream)))
    out = {"faces": r.faces, "surface_measure": r.surface_measure, "flux": r.flux, "traction": list(r.traction),
           "traction_magnitude_integral": r.traction_magnitude, "current": r.current,
           "current_cells": r.current_cells}
    out["mean_traction_magnitude"] = 0.0 if r.surface_measure == 0.0 else r.traction_magnitude / r.surface_measure
    out["terminal_current"] = r.current / num_turns
    out["inductance"] = 0.0 if r.current == 0.0 else r.flux / (r.current * num_turns)
    return out


def ms3d_switches_from_environment():
    """the 3D switches the reference reads from SIMILIE_MAGNETIC_Y_RESPONSE_
    SIGN_3D, SIMILIE_DIVERGENCE_GAUGE_3D and SIMILIE_VECTOR_POTENTIAL_GAUGE_
    SIGN_3D, as keyword arguments of set_switches"""
    sw = L.Ms3dSwitches()
    L.check(L.lib().sil_ms3d_switches_from_environment(ctypes.byref(sw)))
    return {"y_response_sign": sw.y_response_sign, "divergence_gauge": bool(sw.divergence_gauge),
            "gauge_sign": sw.gauge_sign}


class MagnetostaticsOperator2D(LinearOperator):
    """MagnetostaticsOperator2D (magnetostatics_onelab.hpp:1013-1522): A_z on
    a rectilinear node grid, row i + nx·j; interior rows gather ∂H/∂B at the
    interior sampled nodes of the moments B_x = A(s+e_y) − A(s),
    B_y = A(s) − A(s+e_x); boundary rows are the identity.  criterion:
    "moments" (the reference's default), "potential" or
    "potential_and_moments" (row sums counted twice)."""

    def __init__(self, x_coords, y_coords, mu_nodal, criterion="moments"):
        self.shape = (len(x_coords), len(y_coords))
        self.criterion = criterion
        _f64(mu_nodal, "mu_nodal")
        h = ctypes.c_void_p()
        xs, ys = _host_array(x_coords), _host_array(y_coords)
        L.check(L.lib().sil_magnetostatics2d_create(ctypes.byref(h), *self.shape, ctypes.cast(xs, ctypes.c_void_p),
                                                    ctypes.cast(ys, ctypes.c_void_p), L.ptr(mu_nodal),
                                                    CRITERIA[criterion]))
        self.handle = h
        self._destroy = L.lib().sil_operator_destroy

    @staticmethod
    def nodal_mu(shape, cell_mu, mu0, stream=None):
        """mean of the adjacent cells' μ (magnetostatics_onelab.hpp:4467-4500)"""
        nx, ny = shape
        _f64(cell_mu, "cell_mu")
        mu = torch.empty(nx * ny, dtype=torch.float64, device="cuda")
        L.check(L.lib().sil_magnetostatics2d_nodal_mu(nx, ny, L.ptr(cell_mu), float(mu0), L.ptr(mu),
                                                      L.stream_handle(stream)))
        return mu

    def rhs(self, cell_current_z, stream=None):
        """¼·area·J_z over the cells around each interior node
        (magnetostatics_onelab.hpp:4413-4445); cell_current_z: (cells,)"""
        _f64(cell_current_z, "cell_current_z")
        b = torch.empty(self.size(), dtype=torch.float64, device="cuda")
        L.check(L.lib().sil_magnetostatics2d_rhs(self.handle, L.ptr(cell_current_z), L.ptr(b),
                                                 L.stream_handle(stream)))
        return b

    def post_process(self, potential, cell_mu, cell_nonlinear=None, stream=None):
        """CellPostProcessFields of every cell of the 2D grid
        (magnetostatics_onelab.hpp:3866-3982, 4048-4213), cells x fastest"""
        _f64(potential, "potential")
        _f64(cell_mu, "cell_mu")
        if cell_nonlinear is not None:
            _f64(cell_nonlinear, "cell_nonlinear")
        nx, ny = self.shape
        cells = (nx - 1) * (ny - 1)
        out = [torch.empty((cells, w), dtype=torch.float64, device="cuda") for w in (3, 3, 6, 3)]
        L.check(L.lib().sil_magnetostatics2d_post_process(
            self.handle, L.ptr(potential), L.ptr(cell_mu), L.ptr(cell_nonlinear) if cell_nonlinear is not None else None,
            *[L.ptr(o) for o in out], L.stream_handle(stream)))
        return CellPostProcessFields(*out)

    def assemble(self, stream=None):
        """assemble_matrix_data (magnetostatics_onelab.hpp:3554-3748) + the
        CSR conversion (minimize_strong_formulation_residual.hpp:565-646):
        the linear operator as a CSR matrix assembled on the device."""
        h = ctypes.c_void_p()
        L.check(L.lib().sil_magnetostatics2d_assemble(self.handle, ctypes.byref(h), L.stream_handle(stream)))
        return AssembledOperator._adopt(h)


class NonlinearMagnetostaticsOperator2D(MagnetostaticsOperator2D):
    """MagnetostaticsOperator2D with MaterialMagnetostaticsHamiltonian
    (magnetostatics_onelab.hpp:957-1011): apply(x) is F(x) and
    minimize_strong_formulation_residual runs Newton."""

    def __init__(self, x_coords, y_coords, mu_nodal, ferromagnetic_nodal, bh_b, bh_h, criterion="moments"):
        self.shape = (len(x_coords), len(y_coords))
        self.criterion = criterion
        _f64(mu_nodal, "mu_nodal")
        _f64(ferromagnetic_nodal, "ferromagnetic_nodal")
        if len(bh_b) != len(bh_h):
            raise ValueError("bh_b and bh_h must have the same length")
        h = ctypes.c_void_p()
        xs, ys = _host_array(x_coords), _host_array(y_coords)
        bb, hh = _host_array(bh_b), _host_array(bh_h)
        L.check(L.lib().sil_magnetostatics2d_nonlinear_create(
            ctypes.byref(h), *self.shape, ctypes.cast(xs, ctypes.c_void_p), ctypes.cast(ys, ctypes.c_void_p),
            L.ptr(mu_nodal), L.ptr(ferromagnetic_nodal), ctypes.cast(bb, ctypes.c_void_p),
            ctypes.cast(hh, ctypes.c_void_p), len(bh_b), CRITERIA[criterion]))
        self.handle = h
        self._destroy = L.lib().sil_operator_destroy

    @staticmethod
    def nodal_fraction(shape, cell_flags, stream=None):
        """mean over the adjacent cells of a 0/1 nonlinear-material flag"""
        return MagnetostaticsOperator2D.nodal_mu(shape, cell_flags, 0.0, stream)

    def apply_jacobian(self, state, v, out=None, stream=None):
        _f64(state, "state")
        _f64(v, "v")
        out = torch.empty_like(v) if out is None else _f64(out, "out")
        L.check(L.lib().sil_magnetostatics2d_apply_jacobian(self.handle, L.ptr(state), L.ptr(v), L.ptr(out),
                                                             L.stream_handle(stream)))
        return out

    def jacobian_diagonal(self, state, stream=None):
        _f64(state, "state")
        d = torch.empty(self.size(), dtype=torch.float64, device="cuda")
        L.check(L.lib().sil_magnetostatics2d_jacobian_diagonal(self.handle, L.ptr(state), L.ptr(d),
                                                               L.stream_handle(stream)))
        return d

    def assemble_jacobian(self, state, stream=None):
        """assemble_matrix_data(model, state) (magnetostatics_onelab.hpp
        :3554-3748): the Jacobian at `state` as a CSR matrix assembled on the
        device."""
        _f64(state, "state")
        if state.numel() != self.size():
            raise ValueError(f"state must hold {self.size()} values")
        h = ctypes.c_void_p()
        L.check(L.lib().sil_magnetostatics2d_assemble_jacobian(self.handle, L.ptr(state), ctypes.byref(h),
                                                               L.stream_handle(stream)))
        return AssembledOperator._adopt(h)


class Poisson2D(LinearOperator):
    """Dirichlet-zero d0ᵀ ⋆1^κ d0 on the interior of a rectilinear nx×ny grid
    (x slowest); vectors have nx·ny entries with zero boundary entries."""

    def __init__(self, x_coords, y_coords, kappa=None):
        self.shape = (len(x_coords), len(y_coords))
        if kappa is not None:
            _f64(kappa, "kappa")
        self.kappa = kappa
        h = ctypes.c_void_p()
        xs, ys = _host_array(x_coords), _host_array(y_coords)
        L.check(L.lib().sil_poisson2d_create(ctypes.byref(h), self.shape[0], self.shape[1],
                                             ctypes.cast(xs, ctypes.c_void_p), ctypes.cast(ys, ctypes.c_void_p),
                                             L.ptr(kappa)))
        self.handle = h
        self._destroy = L.lib().sil_operator_destroy


@dataclass
class StrongFormulationSolverSettings:
    """minimize_strong_formulation_residual.hpp:119-134 (plus the precision
    mode and the host polling interval of the GPU solver)."""

    max_iterations: int = 2000
    relative_tolerance: float = 1.0e-12
    preconditioner: str = "jacobi"  # "jacobi" | "identity" | "spd_isai"
    precision: str = "fp64"  # "fp64" | "mixed"
    check_interval: int = 0
    replace_interval: int = 0

    def to_c(self):
        s = L.SolverSettings()
        s.max_iterations = int(self.max_iterations)
        s.relative_tolerance = float(self.relative_tolerance)
        s.precision = {"fp64": L.PRECISION_FP64, "mixed": L.PRECISION_MIXED}[self.precision]
        s.preconditioner = {"identity": L.PRECOND_IDENTITY, "jacobi": L.PRECOND_JACOBI,
                            "spd_isai": L.PRECOND_SPD_ISAI}[self.preconditioner]
        s.check_interval = int(self.check_interval)
        s.replace_interval = int(self.replace_interval)
        return s


def minimize_strong_formulation_residual(operator, rhs, solution=None, settings=None, stream=None):
    """x0 = 0, CG until ‖r‖2 ≤ tol·‖b‖2 (minimize_strong_formulation_residual
    .hpp:1308-1382).  Returns (solution, diagnostics dict)."""
    settings = settings or StrongFormulationSolverSettings()
    _f64(rhs, "rhs")
    solution = torch.empty_like(rhs) if solution is None else _f64(solution, "solution")
    diag = L.SolverDiagnostics()
    cs = settings.to_c()
    L.check(L.lib().sil_minimize_strong_formulation_residual(operator.handle, L.ptr(rhs), L.ptr(solution),
                                                             ctypes.byref(cs), ctypes.byref(diag),
                                                             L.stream_handle(stream)))
    return solution, diag.as_dict()
