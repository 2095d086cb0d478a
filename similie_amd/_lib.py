"""Loads libsil_hip.so (the HIP C-ABI library) and declares its signatures.

The library is the only compute path: there is no CPU fallback.  When it
is missing this module raises instead of degrading.
"""
import ctypes
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libsil_hip.so"

SIL_MAX_DIM = 4
SIL_OK = 0
SIL_ERR_INVALID, SIL_ERR_UNSUPPORTED = -1, -2

METRIC_IDENTITY, METRIC_DIAGONAL, METRIC_SYMMETRIC = 0, 1, 2
PRECISION_FP64, PRECISION_MIXED = 0, 1
PRECOND_IDENTITY, PRECOND_JACOBI, PRECOND_LEARNED, PRECOND_SPD_ISAI = 0, 1, 2, 3


class SilError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"libsil_hip status {status}: {message}")
        self.status = status


class View(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int32),
        ("nbatch", ctypes.c_int32),
        ("extent", ctypes.c_int64 * SIL_MAX_DIM),
        ("stride", ctypes.c_int64 * SIL_MAX_DIM),
        ("comp_stride", ctypes.c_int64),
        ("batch_extent", ctypes.c_int64 * SIL_MAX_DIM),
        ("batch_stride", ctypes.c_int64 * SIL_MAX_DIM),
        ("comp_extent", ctypes.c_int64),
    ]


class SolverSettings(ctypes.Structure):
    _fields_ = [
        ("max_iterations", ctypes.c_int32),
        ("relative_tolerance", ctypes.c_double),
        ("precision", ctypes.c_int32),
        ("preconditioner", ctypes.c_int32),
        ("check_interval", ctypes.c_int32),
        ("replace_interval", ctypes.c_int32),
    ]


class SolverDiagnostics(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32),
        ("converged", ctypes.c_int32),
        ("initial_residual_l2", ctypes.c_double),
        ("final_residual_l2", ctypes.c_double),
        ("final_relative_residual", ctypes.c_double),
        ("true_relative_residual", ctypes.c_double),
        ("duration", ctypes.c_double),
        ("outer_iterations", ctypes.c_int32),
        ("verified", ctypes.c_int32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Ms3dSwitches(ctypes.Structure):
    _fields_ = [
        ("y_response_sign", ctypes.c_double),
        ("divergence_gauge", ctypes.c_int32),
        ("gauge_sign", ctypes.c_double),
        ("criterion", ctypes.c_int32),
    ]


class AirGapResult(ctypes.Structure):
    _fields_ = [
        ("faces", ctypes.c_int64),
        ("current_cells", ctypes.c_int64),
        ("surface_measure", ctypes.c_double),
        ("flux", ctypes.c_double),
        ("traction", ctypes.c_double * 3),
        ("traction_magnitude", ctypes.c_double),
        ("current", ctypes.c_double),
    ]


class Lap3dTile(ctypes.Structure):
    _fields_ = [
        ("variant", ctypes.c_int32),
        ("rows", ctypes.c_int32),
        ("xchunk", ctypes.c_int32),
        ("lanes_or_depth", ctypes.c_int32),
        ("nontemporal", ctypes.c_int32),
        ("microseconds", ctypes.c_double),
    ]


class BatchShape(ctypes.Structure):
    _fields_ = [
        ("wide_side", ctypes.c_int32),
        ("wide_workgroups", ctypes.c_int32),
        ("block", ctypes.c_int32),
        ("wide_problems", ctypes.c_int32),
        ("seconds", ctypes.c_double),
    ]


class SilproSettings(ctypes.Structure):
    _fields_ = [
        ("physics", ctypes.c_int32),
        ("max_iterations", ctypes.c_int64),
        ("relative_tolerance", ctypes.c_double),
        ("jacobi_max_block_size", ctypes.c_int32),
        ("use_matrix_free", ctypes.c_int32),
        ("criterion", ctypes.c_int32),
        ("preconditioner", ctypes.c_int32),
        ("vector_potential_gauge_penalty", ctypes.c_double),
        ("sor_relaxation_factor", ctypes.c_double),
        ("chebyshev_iterations", ctypes.c_int32),
        ("chebyshev_lower_bound", ctypes.c_double),
        ("chebyshev_upper_bound", ctypes.c_double),
        ("ir_iterations", ctypes.c_int32),
        ("ir_relaxation_factor", ctypes.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_I32P = ctypes.POINTER(ctypes.c_int32)


class CellMaterials(ctypes.Structure):
    _fields_ = [
        ("mu0", ctypes.c_double), ("core_mu", ctypes.c_double),
        ("current_density_x", ctypes.c_double), ("current_density_z", ctypes.c_double),
        ("use_nonlinear_magnetic_material", ctypes.c_int32),
        ("magnetic_material_tags", _I32P), ("n_magnetic_material_tags", ctypes.c_int32),
        ("positive_x_tags", _I32P), ("n_positive_x_tags", ctypes.c_int32),
        ("negative_x_tags", _I32P), ("n_negative_x_tags", ctypes.c_int32),
        ("positive_z_tags", _I32P), ("n_positive_z_tags", ctypes.c_int32),
        ("negative_z_tags", _I32P), ("n_negative_z_tags", ctypes.c_int32),
    ]


class BatchResult(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32),
        ("converged", ctypes.c_int32),
        ("true_relative_residual", ctypes.c_double),
    ]


class SolveRecord(ctypes.Structure):
    _fields_ = [
        ("every", ctypes.c_int32),
        ("max_records", ctypes.c_int32),
        ("r", ctypes.c_void_p),
        ("x", ctypes.c_void_p),
        ("count", ctypes.POINTER(ctypes.c_int32)),
    ]


_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_D = ctypes.c_double
_VP = ctypes.POINTER(View)

# name: (restype, argtypes)
SIGNATURES = {
    "sil_last_error": (ctypes.c_char_p, []),
    "sil_abi_version": (_I32, []),
    "sil_grid_create": (ctypes.c_int, [ctypes.POINTER(_P), _I32, ctypes.POINTER(_I64), ctypes.POINTER(_P), _P, _I32, _P]),
    "sil_grid_destroy": (ctypes.c_int, [_P]),
    "sil_coboundary": (ctypes.c_int, [_I32, _P, _VP, _P, _VP, _P]),
    "sil_transposed_coboundary": (ctypes.c_int, [_I32, _P, _VP, _P, _VP, _P]),
    "sil_fill_discrete_hodge_star": (ctypes.c_int, [_P, _I32, _P, _P]),
    "sil_hodge_star": (ctypes.c_int, [_P, _I32, _P, _VP, _P, _VP, _P]),
    "sil_codifferential": (ctypes.c_int, [_P, _I32, _P, _VP, _P, _VP, _P]),
    "sil_laplacian": (ctypes.c_int, [_P, _I32, _P, _VP, _P, _VP, _P]),
    "sil_make_staged_laplacian": (ctypes.c_int, [ctypes.POINTER(_P), _P, _I32, _VP, _VP]),
    "sil_staged_laplacian_run": (ctypes.c_int, [_P, _P, _P, _P]),
    "sil_staged_laplacian_destroy": (ctypes.c_int, [_P]),
    "sil_staged_laplacian_autotune": (ctypes.c_int, [_P, _P, _P, _I32, ctypes.POINTER(Lap3dTile), _P]),
    "sil_make_staged_codifferential": (ctypes.c_int, [ctypes.POINTER(_P), _P, _I32, _VP, _VP]),
    "sil_staged_codifferential_run": (ctypes.c_int, [_P, _P, _P, _P]),
    "sil_staged_codifferential_destroy": (ctypes.c_int, [_P]),
    "sil_magnetostatics3d_create": (ctypes.c_int, [ctypes.POINTER(_P), _I64, _I64, _I64, _P, _P, _P, _P, _D]),
    "sil_magnetostatics3d_nodal_mu": (ctypes.c_int, [_I64, _I64, _I64, _P, _D, _P, _P]),
    "sil_magnetostatics3d_rhs": (ctypes.c_int, [_P, _P, _P, _P]),
    "sil_ms3d_switches_from_environment": (ctypes.c_int, [ctypes.POINTER(Ms3dSwitches)]),
    "sil_magnetostatics3d_air_gap_diagnostics": (ctypes.c_int, [_P, _P, _P, _P, _P, ctypes.POINTER(AirGapResult),
                                                                 _P]),
    "sil_magnetostatics2d_air_gap_diagnostics": (ctypes.c_int, [_P, _P, _P, _P, _P, ctypes.c_double,
                                                                 ctypes.c_double, ctypes.POINTER(AirGapResult), _P]),
    "sil_magnetostatics3d_set_switches": (ctypes.c_int, [_P, ctypes.POINTER(Ms3dSwitches)]),
    "sil_magnetostatics3d_post_process": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "sil_magnetostatics2d_create": (ctypes.c_int, [ctypes.POINTER(_P), _I64, _I64, _P, _P, _P, _I32]),
    "sil_magnetostatics2d_nonlinear_create": (
        ctypes.c_int, [ctypes.POINTER(_P), _I64, _I64, _P, _P, _P, _P, _P, _P, _I32, _I32]),
    "sil_magnetostatics2d_apply_jacobian": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "sil_magnetostatics2d_jacobian_diagonal": (ctypes.c_int, [_P, _P, _P, _P]),
    "sil_magnetostatics2d_nodal_mu": (ctypes.c_int, [_I64, _I64, _P, _D, _P, _P]),
    "sil_magnetostatics2d_rhs": (ctypes.c_int, [_P, _P, _P, _P]),
    "sil_magnetostatics2d_post_process": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "sil_poisson2d_create": (ctypes.c_int, [ctypes.POINTER(_P), _I64, _I64, _P, _P, _P]),
    "sil_magnetostatics3d_nonlinear_create": (
        ctypes.c_int, [ctypes.POINTER(_P), _I64, _I64, _I64, _P, _P, _P, _P, _P, _P, _P, _I32, _D]),
    "sil_magnetostatics3d_apply_jacobian": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "sil_magnetostatics3d_jacobian_diagonal": (ctypes.c_int, [_P, _P, _P, _P]),
    "sil_magnetostatics3d_assemble": (ctypes.c_int, [_P, ctypes.POINTER(_P), _P]),
    "sil_magnetostatics3d_assemble_jacobian": (ctypes.c_int, [_P, _P, ctypes.POINTER(_P), _P]),
    "sil_magnetostatics2d_assemble": (ctypes.c_int, [_P, ctypes.POINTER(_P), _P]),
    "sil_magnetostatics2d_assemble_jacobian": (ctypes.c_int, [_P, _P, ctypes.POINTER(_P), _P]),
    "sil_csr_create": (ctypes.c_int, [ctypes.POINTER(_P), _I64, _I64, _P, _P, _P, _P]),
    "sil_csr_nnz": (_I64, [_P]),
    "sil_csr_copy": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "sil_spd_isai_create": (ctypes.c_int, [_P, ctypes.POINTER(_P), _P]),
    "sil_msh2_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_P)]),
    "sil_msh2_parse": (ctypes.c_int, [ctypes.c_char_p, _I64, ctypes.POINTER(_P)]),
    "sil_structured_grid_shape": (ctypes.c_int, [_P, ctypes.POINTER(_I32), ctypes.POINTER(_I64)]),
    "sil_structured_grid_coords": (ctypes.c_int, [_P, _I32, _P]),
    "sil_structured_grid_plane": (ctypes.c_double, [_P]),
    "sil_structured_grid_cell_tags": (ctypes.c_int, [_P, _P]),
    "sil_structured_grid_destroy": (ctypes.c_int, [_P]),
    "sil_silpro_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_P)]),
    "sil_silpro_parse": (ctypes.c_int, [ctypes.c_char_p, _I64, ctypes.POINTER(_P)]),
    "sil_silpro_value": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_char_p, _I64]),
    "sil_silpro_section_value": (ctypes.c_int, [_P, ctypes.c_char_p, _I32, ctypes.c_char_p, _I64]),
    "sil_silpro_problem_settings": (ctypes.c_int, [_P, ctypes.POINTER(SilproSettings)]),
    "sil_silpro_solver_settings": (ctypes.c_int, [ctypes.POINTER(SilproSettings), ctypes.POINTER(SolverSettings)]),
    "sil_silpro_destroy": (ctypes.c_int, [_P]),
    "sil_magnetostatics_cell_inputs": (ctypes.c_int, [_P, _I64, ctypes.POINTER(CellMaterials), _P, _P, _P]),
    "sil_operator_size": (_I64, [_P]),
    "sil_operator_apply": (ctypes.c_int, [_P, _P, _P, _P]),
    "sil_operator_apply_f32": (ctypes.c_int, [_P, _P, _P, _P]),
    "sil_operator_diagonal": (ctypes.c_int, [_P, _P, _P]),
    "sil_operator_destroy": (ctypes.c_int, [_P]),
    "sil_default_solver_settings": (None, [ctypes.POINTER(SolverSettings)]),
    "sil_minimize_strong_formulation_residual": (
        ctypes.c_int,
        [_P, _P, _P, ctypes.POINTER(SolverSettings), ctypes.POINTER(SolverDiagnostics), _P],
    ),
    "sil_minimize_strong_formulation_residual_noalloc": (
        ctypes.c_int,
        [_P, _P, _P, ctypes.POINTER(SolverSettings), ctypes.POINTER(SolverDiagnostics), _P],
    ),
    "sil_noalloc_max_unknowns": (_I64, []),
    "sil_batched_poisson2d_create": (ctypes.c_int, [ctypes.POINTER(_P), _I32, ctypes.POINTER(_I32), _P]),
    "sil_batched_poisson2d_offset": (_I64, [_P, _I32]),
    "sil_batched_poisson2d_set_block": (ctypes.c_int, [_P, _I32]),
    "sil_batched_poisson2d_autotune": (
        ctypes.c_int,
        [_P, _P, _I32, ctypes.POINTER(_I32), ctypes.POINTER(ctypes.c_double)],
    ),
    "sil_batched_poisson2d_autotune_shape": (ctypes.c_int, [_P, _P, _I32, ctypes.POINTER(BatchShape)]),
    "sil_batched_poisson2d_solve": (
        ctypes.c_int,
        [_P, _P, _P, ctypes.POINTER(SolverSettings), ctypes.POINTER(BatchResult), ctypes.POINTER(ctypes.c_double), _P],
    ),
    "sil_batched_poisson2d_destroy": (ctypes.c_int, [_P]),
    "sil_unet_param_count": (_I64, []),
    "sil_unet_create": (ctypes.c_int, [ctypes.POINTER(_P), _I32, _I32, _I32, ctypes.c_uint64]),
    "sil_unet_destroy": (ctypes.c_int, [_P]),
    "sil_unet_get_params": (ctypes.c_int, [_P, _P, _P]),
    "sil_unet_set_params": (ctypes.c_int, [_P, _P, _P]),
    "sil_unet_activation": (ctypes.c_int, [_P, _I32, _I32, _P, _P]),
    "sil_unet_forward": (ctypes.c_int, [_P, _I32, _P, _P, _P]),
    "sil_unet_infer": (ctypes.c_int, [_P, _I32, _P, _P, _P]),
    "sil_unet_backward": (ctypes.c_int, [_P, _I32, _P, _P, _P, _P]),
    "sil_unet_loss": (ctypes.c_int, [_P, _I32, _P, _P, _P, _P, _P, _P, _P, _P]),
    "sil_unet_train_step": (ctypes.c_int, [_P, _I32, _P, _P, _P, _P, ctypes.c_float, _P, _P]),
    "sil_poisson2d_batch_create": (ctypes.c_int, [ctypes.POINTER(_P), _I32, _I32, _P]),
    "sil_poisson2d_batch_destroy": (ctypes.c_int, [_P]),
    "sil_poisson2d_batch_apply": (ctypes.c_int, [_P, _P, _P, _P]),
    "sil_poisson2d_batch_weights": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "sil_poisson2d_batch_solve": (
        ctypes.c_int,
        [_P, _P, _P, ctypes.POINTER(SolverSettings), _P, ctypes.POINTER(BatchResult),
         ctypes.POINTER(ctypes.c_double), ctypes.POINTER(SolveRecord), _P],
    ),
}

_lib = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m similie_amd.build` "
                "(there is no CPU fallback for the HIP operators)"
            )
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status):
    if status != SIL_OK:
        msg = lib().sil_last_error()
        raise SilError(status, msg.decode() if msg else "")
    return status


def stream_handle(stream=None):
    """HIP stream handle of a torch stream (torch's current stream by default)."""
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None
