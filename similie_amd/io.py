"""Input formats: Gmsh MSH2 structured grids, .silpro problem files and the
cell inputs of the magnetostatics preprocess.

Python mirror of the reference's ONELAB input side
(onelab_interface/gmsh_structured_grid.hpp:239-553,
onelab_interface/onelab_interface.hpp:160-560,
onelab_interface/magnetostatics_onelab.hpp:4870-4897).  The parsing runs in
libsil_hip.so on the host; nothing here touches the GPU.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib as L

PHYSICS = {0: "LinearMagnetostatics", 1: "NonLinearMagnetostatics", 2: "ScalarFieldWithPowerCoupling"}
PRECONDITIONERS = {0: "Identity", 1: "Jacobi", 2: "SpdIsai", 3: "SymmetricGaussSeidel", 4: "Ssor",
                   5: "ChebyshevJacobi", 6: "IrJacobi", 7: "GeneralIsai"}
CRITERIA = {0: "PotentialTemporalDerivative", 1: "MomentsTemporalDerivative",
            2: "PotentialAndMomentsTemporalDerivative"}


@dataclass
class StructuredGrid:
    """A rectilinear grid read from a mesh: axis coordinates (x, y[, z]),
    the physical tag of every cell (x fastest) and, for a quadrilateral
    mesh, the z of its plane."""

    coords: list
    cell_tags: np.ndarray
    plane: float = 0.0

    @property
    def ndim(self):
        return len(self.coords)

    @property
    def shape(self):
        return tuple(len(c) for c in self.coords)

    @property
    def cell_shape(self):
        return tuple(len(c) - 1 for c in self.coords)


def _grid_from_handle(h):
    try:
        nd = ctypes.c_int32()
        nodes = (ctypes.c_int64 * 3)()
        L.check(L.lib().sil_structured_grid_shape(h, ctypes.byref(nd), nodes))
        coords = []
        for a in range(nd.value):
            c = np.empty(nodes[a], dtype=np.float64)
            L.check(L.lib().sil_structured_grid_coords(h, a, c.ctypes.data_as(ctypes.c_void_p)))
            coords.append(c)
        ncells = int(np.prod([n - 1 for n in nodes[: nd.value]]))
        tags = np.empty(ncells, dtype=np.int32)
        L.check(L.lib().sil_structured_grid_cell_tags(h, tags.ctypes.data_as(ctypes.c_void_p)))
        plane = float(L.lib().sil_structured_grid_plane(h))
        return StructuredGrid(coords, tags, plane)
    finally:
        L.lib().sil_structured_grid_destroy(h)


def read_msh2(path):
    """gmsh::parse_supported_msh2_mesh + build_structured_grid of a file"""
    h = ctypes.c_void_p()
    L.check(L.lib().sil_msh2_read(str(path).encode(), ctypes.byref(h)))
    return _grid_from_handle(h)


def parse_msh2(text):
    """the same from the file's text (str or bytes)"""
    data = text.encode() if isinstance(text, str) else bytes(text)
    h = ctypes.c_void_p()
    L.check(L.lib().sil_msh2_parse(data, len(data), ctypes.byref(h)))
    return _grid_from_handle(h)


class Silpro:
    """A parsed .silpro file: values by dotted path, section value lists, and
    the Problem / Solver settings with the reference's defaults."""

    def __init__(self, handle):
        self.handle = handle
        self._destroy = L.lib().sil_silpro_destroy

    @classmethod
    def read(cls, path):
        h = ctypes.c_void_p()
        L.check(L.lib().sil_silpro_read(str(path).encode(), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def parse(cls, text):
        data = text.encode() if isinstance(text, str) else bytes(text)
        h = ctypes.c_void_p()
        L.check(L.lib().sil_silpro_parse(data, len(data), ctypes.byref(h)))
        return cls(h)

    def value(self, path, default=None):
        buf = ctypes.create_string_buffer(4096)
        rc = L.lib().sil_silpro_value(self.handle, path.encode(), buf, len(buf))
        if rc != L.SIL_OK:
            if default is not None:
                return default
            L.check(rc)
        return buf.value.decode()

    def values(self, section):
        """the values of a section in key order ([] when it is absent)"""
        out = []
        buf = ctypes.create_string_buffer(4096)
        while L.lib().sil_silpro_section_value(self.handle, section.encode(), len(out), buf, len(buf)) == L.SIL_OK:
            out.append(buf.value.decode())
        return out

    def problem_settings(self):
        st = L.SilproSettings()
        L.check(L.lib().sil_silpro_problem_settings(self.handle, ctypes.byref(st)))
        d = st.as_dict()
        d["physics"] = PHYSICS[d["physics"]]
        d["preconditioner"] = PRECONDITIONERS[d["preconditioner"]]
        d["criterion"] = CRITERIA[d["criterion"]]
        d["use_matrix_free"] = bool(d["use_matrix_free"])
        return d

    def solver_settings(self, precision="fp64"):
        """StrongFormulationSolverSettings of this library for the problem"""
        from .solvers import StrongFormulationSolverSettings

        st = L.SilproSettings()
        L.check(L.lib().sil_silpro_problem_settings(self.handle, ctypes.byref(st)))
        out = L.SolverSettings()
        L.check(L.lib().sil_silpro_solver_settings(ctypes.byref(st), ctypes.byref(out)))
        names = {L.PRECOND_IDENTITY: "identity", L.PRECOND_JACOBI: "jacobi", L.PRECOND_SPD_ISAI: "spd_isai"}
        return StrongFormulationSolverSettings(max_iterations=out.max_iterations,
                                               relative_tolerance=out.relative_tolerance,
                                               preconditioner=names[out.preconditioner], precision=precision)

    def __del__(self):
        h = getattr(self, "handle", None)
        fn = getattr(self, "_destroy", None)
        if h is not None and h.value and fn is not None:
            fn(h)
            self.handle = None


def _tags(seq):
    arr = np.ascontiguousarray(list(seq), dtype=np.int32)
    return arr, arr.ctypes.data_as(L._I32P), len(arr)


def magnetostatics_cell_inputs(cell_tags, mu0, core_mu=0.0, current_density_x=0.0, current_density_z=0.0,
                               magnetic_material_tags=(), positive_x_tags=(), negative_x_tags=(),
                               positive_z_tags=(), negative_z_tags=(), nonlinear=False):
    """(cell μ, cell current [cells][3], nonlinear-material flags) from the
    physical tags (magnetostatics_onelab.hpp:4870-4897)"""
    tags = np.ascontiguousarray(cell_tags, dtype=np.int32)
    n = tags.size
    keep = []
    m = L.CellMaterials()
    m.mu0, m.core_mu = float(mu0), float(core_mu)
    m.current_density_x, m.current_density_z = float(current_density_x), float(current_density_z)
    m.use_nonlinear_magnetic_material = 1 if nonlinear else 0
    for field, seq in (("magnetic_material_tags", magnetic_material_tags), ("positive_x_tags", positive_x_tags),
                       ("negative_x_tags", negative_x_tags), ("positive_z_tags", positive_z_tags),
                       ("negative_z_tags", negative_z_tags)):
        arr, ptr, cnt = _tags(seq)
        keep.append(arr)
        setattr(m, field, ptr)
        setattr(m, "n_" + field, cnt)
    mu = np.empty(n)
    current = np.empty((n, 3))
    flags = np.empty(n)
    L.check(L.lib().sil_magnetostatics_cell_inputs(tags.ctypes.data_as(ctypes.c_void_p), n, ctypes.byref(m),
                                                   mu.ctypes.data_as(ctypes.c_void_p),
                                                   current.ctypes.data_as(ctypes.c_void_p),
                                                   flags.ctypes.data_as(ctypes.c_void_p)))
    return mu, current, flags
