"""ctypes access to the CPU oracle (oracle/liboracle.so) — the checker.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use
this module.  Arrays are numpy float64 in the reference's index-last layout
(axis 0 slowest, component fastest).
"""
import ctypes
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
_LIB_PATH = ROOT / "oracle" / "liboracle.so"
_lib = None

POS_RECTILINEAR, POS_NODAL = 0, 1
METRIC_IDENTITY, METRIC_DIAGONAL, METRIC_SYMMETRIC = 0, 1, 2

_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_long)


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
        _lib = ctypes.CDLL(str(_LIB_PATH))
        _lib.orc_binom.restype = ctypes.c_long
        _lib.orc_hodge_coefficient.restype = ctypes.c_double
        _lib.orc_hodge_coefficient_chain.restype = ctypes.c_double
        _lib.orc_continuous_hodge_value.restype = ctypes.c_double
        _lib.orc_cg.restype = ctypes.c_int
        _lib.orc_poisson2d_cg.restype = ctypes.c_int
        _lib.orc_ms_cg.restype = ctypes.c_int
        _lib.orc_ms_assemble_csr.restype = ctypes.c_long
        _lib.orc_spd_isai.restype = ctypes.c_long
        _lib.orc_bh_eval.restype = ctypes.c_double
        _lib.orc_nl_newton.restype = ctypes.c_int
        _lib.orc_ms2d_newton.restype = ctypes.c_int
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _l(seq):
    arr = (ctypes.c_long * len(seq))(*[int(v) for v in seq])
    return arr


def binom(n, k):
    return int(lib().orc_binom(n, k))


def combo_mask(n, k, index):
    return int(lib().orc_combo_mask(n, k, index))


def combo_index(n, k, mask):
    return int(lib().orc_combo_index(n, k, mask))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Geometry:
    """Grid geometry for the oracle: rectilinear axis coordinates or nodal
    positions, plus an identity / diagonal / symmetric metric per node."""

    def __init__(self, ext, axes=None, nodal=None, metric_kind=METRIC_IDENTITY, metric=None):
        self.ext = [int(e) for e in ext]
        self.ndim = len(self.ext)
        self.pos_kind = POS_NODAL if nodal is not None else POS_RECTILINEAR
        self.axes = [_f64(a) for a in axes] if axes is not None else None
        self.nodal = _f64(nodal) if nodal is not None else None
        self.metric_kind = metric_kind
        self.metric = _f64(metric) if metric is not None else None
        self._keep = []

    def args(self):
        n = self.ndim
        ext = _l(self.ext)
        if self.axes is not None:
            ax = (_dp * 4)(*([_d(a) for a in self.axes] + [None] * (4 - n)))
        else:
            ax = None
        nodal = _d(self.nodal) if self.nodal is not None else None
        metric = _d(self.metric) if self.metric is not None else None
        self._keep = [ext, ax]
        return [ctypes.c_int(n), ext, ctypes.c_int(self.pos_kind), ax, nodal, ctypes.c_int(self.metric_kind), metric]


def coboundary(n, k, in_ext, arr, out_ext):
    nout = binom(n, k + 1)
    out = np.zeros(tuple(out_ext) + (nout,))
    a = _f64(arr)
    lib().orc_coboundary(n, k, _l(in_ext), _d(a), _l(out_ext), _d(out))
    return out


def transposed_coboundary(n, k, in_ext, arr, out_ext):
    nout = binom(n, k + 1)
    out = np.zeros(tuple(out_ext) + (nout,))
    a = _f64(arr)
    lib().orc_transposed_coboundary(n, k, _l(in_ext), _d(a), _l(out_ext), _d(out))
    return out


def hodge_star(geom, k, arr):
    n = geom.ndim
    out = np.zeros(tuple(geom.ext) + (binom(n, n - k),))
    a = _f64(arr)
    lib().orc_hodge_star(*geom.args(), ctypes.c_int(k), _d(a), _d(out))
    return out


def fill_hodge_star(geom, k, use_chain=False):
    n = geom.ndim
    out = np.zeros(tuple(geom.ext) + (n ** k, binom(n, n - k)))
    lib().orc_fill_hodge_star(*geom.args(), ctypes.c_int(k), ctypes.c_int(1 if use_chain else 0), _d(out))
    return out


def codifferential(geom, k, arr, out_ext):
    n = geom.ndim
    out = np.zeros(tuple(out_ext) + (binom(n, k - 1),))
    a = _f64(arr)
    lib().orc_codifferential(*geom.args(), ctypes.c_int(k), _d(a), _l(out_ext), _d(out))
    return out


def laplacian(geom, k, arr, out_ext):
    n = geom.ndim
    out = np.zeros(tuple(out_ext) + (binom(n, k),))
    a = _f64(arr)
    lib().orc_laplacian(*geom.args(), ctypes.c_int(k), _d(a), _l(out_ext), _d(out))
    return out


def reduction(geom, k, arr):
    n = geom.ndim
    out = np.zeros(tuple(geom.ext) + (binom(n, k),))
    a = _f64(arr)
    args = geom.args()
    lib().orc_reduction(args[0], args[1], args[2], args[3], args[4], ctypes.c_int(k), _d(a), _d(out))
    return out


def reconstruction(geom, k, arr):
    n = geom.ndim
    out = np.zeros(tuple(geom.ext) + (binom(n, k),))
    a = _f64(arr)
    args = geom.args()
    lib().orc_reconstruction(args[0], args[1], args[2], args[3], args[4], ctypes.c_int(k), _d(a), _d(out))
    return out


def hodge_coefficient(geom, node, src_mask, tgt_mask, chain=False):
    """Coefficient of ⋆ from the canonical source axes set to the target set
    at one node, by the volume ratio or by the reference's full chain."""
    import ctypes as C

    class _G(C.Structure):
        _fields_ = [
            ("ndim", C.c_int),
            ("ext", C.c_long * 4),
            ("pos_kind", C.c_int),
            ("axis", _dp * 4),
            ("nodal", _dp),
            ("metric_kind", C.c_int),
            ("metric", _dp),
        ]

    g = _G()
    g.ndim = geom.ndim
    for a in range(geom.ndim):
        g.ext[a] = geom.ext[a]
        if geom.axes is not None:
            g.axis[a] = _d(geom.axes[a])
    g.pos_kind = geom.pos_kind
    g.nodal = _d(geom.nodal) if geom.nodal is not None else None
    g.metric_kind = geom.metric_kind
    g.metric = _d(geom.metric) if geom.metric is not None else None
    fn = lib().orc_hodge_coefficient_chain if chain else lib().orc_hodge_coefficient
    return float(fn(C.byref(g), _l(node), C.c_uint(src_mask), C.c_uint(tgt_mask)))


def continuous_hodge_value(ext, metric_kind, metric, node, src, tgt):
    m = _f64(metric)
    return float(lib().orc_continuous_hodge_value(len(ext), _l(ext), metric_kind, _d(m), _l(node), src, tgt))


def boundary(n, vmask):
    k1 = bin(vmask).count("1")
    offs = (ctypes.c_long * (2 * k1 * n))()
    vms = (ctypes.c_int * (2 * k1))()
    negs = (ctypes.c_int * (2 * k1))()
    nf = lib().orc_boundary(n, vmask, offs, vms, negs)
    return [([offs[f * n + a] for a in range(n)], vms[f], negs[f]) for f in range(nf)]


# ---------------------------------------------------------------- solvers ---

def ms_apply(coords, mu, penalty, x, y_sign=1.0):
    nx, ny, nz = (len(c) for c in coords)
    xc, yc, zc = (_f64(c) for c in coords)
    mu = _f64(mu)
    xin = _f64(x)
    out = np.zeros(3 * nx * ny * nz)
    lib().orc_ms_apply(ctypes.c_long(nx), ctypes.c_long(ny), ctypes.c_long(nz), _d(xc), _d(yc), _d(zc), _d(mu),
                       ctypes.c_double(penalty), ctypes.c_double(y_sign), _d(xin), _d(out))
    return out


class ms_switches:
    """the 3D operator's switches for every oracle 3D function inside the
    block: sign of the y moment's response, divergence gauge, gauge sign"""

    def __init__(self, y_sign=1.0, divergence_gauge=False, gauge_sign=1.0):
        self.args = (float(y_sign), int(bool(divergence_gauge)), float(gauge_sign))

    def __enter__(self):
        lib().orc_ms_set_switches(ctypes.c_double(self.args[0]), ctypes.c_int(self.args[1]),
                                  ctypes.c_double(self.args[2]))
        return self

    def __exit__(self, *exc):
        lib().orc_ms_set_switches(ctypes.c_double(1.0), ctypes.c_int(0), ctypes.c_double(1.0))
        return False


def ms_codifferential(coords, x):
    """δA per node of the divergence gauge's codifferential (interior DOFs)"""
    nx, ny, nz = (len(c) for c in coords)
    xc, yc, zc = (_f64(c) for c in coords)
    out = np.zeros(nx * ny * nz)
    xin = _f64(x)
    lib().orc_ms_codifferential(ctypes.c_long(nx), ctypes.c_long(ny), ctypes.c_long(nz), _d(xc), _d(yc), _d(zc),
                                _d(xin), _d(out))
    return out


def ms_rhs(coords, cell_current):
    nx, ny, nz = (len(c) for c in coords)
    xc, yc, zc = (_f64(c) for c in coords)
    cj = _f64(cell_current)
    out = np.zeros(3 * nx * ny * nz)
    lib().orc_ms_rhs(ctypes.c_long(nx), ctypes.c_long(ny), ctypes.c_long(nz), _d(xc), _d(yc), _d(zc), _d(cj), _d(out))
    return out


def ms_nodal_mu(shape, cell_mu, mu0):
    nx, ny, nz = shape
    cm = _f64(cell_mu)
    out = np.zeros(nx * ny * nz)
    lib().orc_ms_nodal_mu(ctypes.c_long(nx), ctypes.c_long(ny), ctypes.c_long(nz), _d(cm), ctypes.c_double(mu0), _d(out))
    return out


def ms_diagonal(coords, mu, penalty):
    nx, ny, nz = (len(c) for c in coords)
    xc, yc, zc = (_f64(c) for c in coords)
    mu = _f64(mu)
    out = np.zeros(3 * nx * ny * nz)
    lib().orc_ms_diagonal(ctypes.c_long(nx), ctypes.c_long(ny), ctypes.c_long(nz), _d(xc), _d(yc), _d(zc), _d(mu),
                          ctypes.c_double(penalty), _d(out))
    return out


def ms_cg(coords, mu, penalty, b, tol, max_it, diag=None):
    nx, ny, nz = (len(c) for c in coords)
    xc, yc, zc = (_f64(c) for c in coords)
    mu = _f64(mu)
    bb = _f64(b)
    x = np.zeros_like(bb)
    res = np.zeros(2)
    dg = _d(_f64(diag)) if diag is not None else None
    it = lib().orc_ms_cg(ctypes.c_long(nx), ctypes.c_long(ny), ctypes.c_long(nz), _d(xc), _d(yc), _d(zc), _d(mu),
                         ctypes.c_double(penalty), _d(bb), dg, ctypes.c_double(tol), ctypes.c_int(max_it), _d(x),
                         _d(res))
    return x, it, res[0], res[1]


def ms_assemble_csr(coords, mu, penalty, y_sign=1.0):
    """(row_ptr int64, cols int32, vals float64) of the assembled operator,
    ascending columns per row."""
    nx, ny, nz = (len(c) for c in coords)
    xc, yc, zc = (_f64(c) for c in coords)
    mu = _f64(mu)
    n = 3 * nx * ny * nz
    rp = np.zeros(n + 1, dtype=np.int64)
    cols = np.zeros(17 * n, dtype=np.int32)
    vals = np.zeros(17 * n)
    nnz = lib().orc_ms_assemble_csr(ctypes.c_long(nx), ctypes.c_long(ny), ctypes.c_long(nz), _d(xc), _d(yc),
                                    _d(zc), _d(mu), ctypes.c_double(penalty), ctypes.c_double(y_sign),
                                    rp.ctypes.data_as(_lp), cols.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                    _d(vals))
    if nnz < 0:
        raise RuntimeError("row capacity exceeded")
    return rp, cols[:nnz].copy(), vals[:nnz].copy()


def spd_isai(row_ptr, cols, vals):
    """factor G of the SPD sparse approximate inverse (M = Gᵀ G), CSR with
    the pattern of tril(A)"""
    n = len(row_ptr) - 1
    rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
    cc = np.ascontiguousarray(cols, dtype=np.int32)
    vv = _f64(vals)
    grp = np.zeros(n + 1, dtype=np.int64)
    gc = np.zeros(max(len(cc), 1), dtype=np.int32)
    gv = np.zeros(max(len(cc), 1))
    rc = lib().orc_spd_isai(ctypes.c_long(n), rp.ctypes.data_as(_lp), cc.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                            _d(vv), grp.ctypes.data_as(_lp), gc.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _d(gv))
    if rc != 0:
        raise ValueError(f"spd_isai failed with code {rc}")
    nnz = int(grp[-1])
    return grp, gc[:nnz].copy(), gv[:nnz].copy()


# --------------------------------------------------- nonlinear magnetostatics ---

BH_FUNCTIONS = {"h_from_b": 0, "dh_db_from_b": 1, "h_over_b_from_q": 2, "d_h_over_b_dq_from_q": 3,
                "d2_h_over_b_dq2_from_q": 4, "nu": 5}


def bh_eval(bh_b, bh_h, name, arg):
    b, h = _f64(bh_b), _f64(bh_h)
    return float(lib().orc_bh_eval(_d(b), _d(h), ctypes.c_int(len(b)), ctypes.c_int(BH_FUNCTIONS[name]),
                                   ctypes.c_double(arg)))


class Nonlinear:
    """grid, nodal μ, ferromagnetic fraction, B-H samples and gauge penalty
    of the nonlinear magnetostatics operator"""

    def __init__(self, coords, mu, ferro, bh_b, bh_h, penalty):
        self.shape = tuple(len(c) for c in coords)
        self.coords = [_f64(c) for c in coords]
        self.mu, self.ferro = _f64(mu), _f64(ferro)
        self.bh_b, self.bh_h = _f64(bh_b), _f64(bh_h)
        self.penalty = float(penalty)
        self.n = 3 * int(np.prod(self.shape))

    def _head(self):
        nx, ny, nz = self.shape
        xc, yc, zc = self.coords
        return [ctypes.c_long(nx), ctypes.c_long(ny), ctypes.c_long(nz), _d(xc), _d(yc), _d(zc), _d(self.mu),
                _d(self.ferro), _d(self.bh_b), _d(self.bh_h), ctypes.c_int(len(self.bh_b)),
                ctypes.c_double(self.penalty)]

    def apply(self, x):
        out = np.zeros(self.n)
        xx = _f64(x)
        lib().orc_nl_apply(*self._head(), _d(xx), _d(out))
        return out

    def jacobian_apply(self, state, v):
        out = np.zeros(self.n)
        st, vv = _f64(state), _f64(v)
        lib().orc_nl_jacobian_apply(*self._head(), _d(st), _d(vv), _d(out))
        return out

    def jacobian_diagonal(self, state):
        out = np.zeros(self.n)
        st = _f64(state)
        lib().orc_nl_jacobian_diagonal(*self._head(), _d(st), _d(out))
        return out

    def jacobian_csr(self, state):
        """(row_ptr int64, cols int32, vals float64) of J(state) assembled,
        ascending columns per row"""
        rp = np.zeros(self.n + 1, dtype=np.int64)
        cols = np.zeros(27 * self.n, dtype=np.int32)
        vals = np.zeros(27 * self.n)
        st = _f64(state)
        lib().orc_nl_jacobian_assemble_csr.restype = ctypes.c_long
        nnz = lib().orc_nl_jacobian_assemble_csr(*self._head(), _d(st), rp.ctypes.data_as(_lp),
                                                 cols.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _d(vals))
        if nnz < 0:
            raise RuntimeError("row capacity exceeded")
        return rp, cols[:nnz].copy(), vals[:nnz].copy()

    def newton(self, b, tol, max_inner, jacobi=True):
        x = np.zeros(self.n)
        out = np.zeros(4)
        bb = _f64(b)
        # jacobi: True (Jacobi), False (none) or "spd_isai" (SpdIsai of the
        # Jacobian assembled at every step)
        pc = 2 if jacobi == "spd_isai" else (1 if jacobi else 0)
        ok = lib().orc_nl_newton(*self._head(), _d(bb), ctypes.c_double(tol), ctypes.c_int(max_inner),
                                 ctypes.c_int(pc), _d(x), _d(out))
        return x, {"converged": bool(ok), "newton_steps": int(out[0]), "iterations": int(out[1]),
                   "initial_residual_l2": out[2], "final_residual_l2": out[3]}


def ms_post_process(coords, potential, cell_mu, cell_nonlinear=None, bh_b=None, bh_h=None):
    """cell induction [cells][3], field [cells][3], Maxwell stress [cells][6]
    and force density [cells][3] of a 3D potential [nodes][3] (cells x
    fastest)"""
    nx, ny, nz = (len(c) for c in coords)
    xc, yc, zc = (_f64(c) for c in coords)
    ncell = (nx - 1) * (ny - 1) * (nz - 1)
    pot, mu = _f64(potential), _f64(cell_mu)
    nl = _f64(cell_nonlinear) if cell_nonlinear is not None else None
    bb = _f64(bh_b if bh_b is not None else [0.0])
    hh = _f64(bh_h if bh_h is not None else [0.0])
    out = [np.zeros((ncell, w)) for w in (3, 3, 6, 3)]
    lib().orc_ms_post_process(ctypes.c_long(nx), ctypes.c_long(ny), ctypes.c_long(nz), _d(xc), _d(yc), _d(zc),
                              _d(pot), _d(mu), _d(nl) if nl is not None else None, _d(bb), _d(hh),
                              ctypes.c_int(len(bb) if bh_b is not None else 0), *[_d(o) for o in out])
    return tuple(out)


def poisson2d_apply(hx, hy, kappa, u):
    nx, ny = u.shape
    hx, hy, uu = _f64(hx), _f64(hy), _f64(u)
    kp = _d(_f64(kappa)) if kappa is not None else None
    out = np.zeros((nx, ny))
    lib().orc_poisson2d_apply(ctypes.c_long(nx), ctypes.c_long(ny), _d(hx), _d(hy), kp, _d(uu), _d(out))
    return out


def poisson2d_cg(hx, hy, kappa, b, tol, max_it):
    nx, ny = b.shape
    hx, hy, bb = _f64(hx), _f64(hy), _f64(b)
    kp = _d(_f64(kappa)) if kappa is not None else None
    x = np.zeros((nx, ny))
    res = np.zeros(2)
    it = lib().orc_poisson2d_cg(ctypes.c_long(nx), ctypes.c_long(ny), _d(hx), _d(hy), kp, _d(bb),
                                ctypes.c_double(tol), ctypes.c_int(max_it), _d(x), _d(res))
    return x, it, res[0], res[1]


# ------------------------------------------------------- CPU baseline ---

class StagedLaplacian3D1:
    """Reference-structured CPU staged Laplacian (90 ⋆ doubles per node, six
    OpenMP passes) used as bench.py's CPU baseline; `L` is the library it
    runs from (the literal oracle build by default)."""

    def __init__(self, ext, axes, metric_diag, L=None):
        L = L or lib()
        self.L = L
        L.orc_staged_lap3d1_create.restype = ctypes.c_void_p
        L.orc_staged_lap3d1_run.argtypes = [ctypes.c_void_p, _dp, _dp, _lp]
        L.orc_staged_lap3d1_destroy.argtypes = [ctypes.c_void_p]
        self.ext = [int(e) for e in ext]
        self.axes = [_f64(a) for a in axes]
        self.metric = _f64(metric_diag) if metric_diag is not None else None
        ax = (_dp * 3)(*[_d(a) for a in self.axes])
        self.h = L.orc_staged_lap3d1_create(_l(self.ext), ax, _d(self.metric) if self.metric is not None else None)

    def run(self, w, out_ext, out=None):
        if out is None:
            out = np.zeros(tuple(out_ext) + (3,))
        ww = _f64(w)
        self.L.orc_staged_lap3d1_run(self.h, _d(ww), _d(out), _l(out_ext))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_staged_lap3d1_destroy(self.h)
            self.h = None


class CpuBaseline:
    """The CPU baselines of oracle/cpu_baseline.c and oracle/cpu_fast.c from
    one build: the literal oracle library, or the -O3 -march=native build
    that bench.py makes on the host it times (build_native)."""

    def __init__(self, path=None):
        self.L = ctypes.CDLL(str(path)) if path else lib()
        self.L.orc_threads.restype = ctypes.c_int
        self.L.orc_fast_poisson2d_cg.restype = ctypes.c_int
        self.L.orc_fast_ms3d_cg.restype = ctypes.c_int

    def threads(self):
        return int(self.L.orc_threads())

    def staged(self, ext, axes, metric_diag):
        return StagedLaplacian3D1(ext, axes, metric_diag, L=self.L)

    def fused_lap3d1(self, ext, axes, metric_diag, w, out_ext, out=None):
        """Δ of a 1-form, one sweep per thread; metric_diag [nodes][3]"""
        if out is None:
            out = np.zeros(tuple(out_ext) + (3,))
        axes = [_f64(a) for a in axes]
        ax = (_dp * 3)(*[_d(a) for a in axes])
        g, ww = _f64(metric_diag), _f64(w)
        self.L.orc_fused_lap3d1(_l(ext), ax, _d(g), _d(ww), _d(out), _l(out_ext))
        return out

    def poisson2d_cg(self, hx, hy, kappa, b, tol, max_it, jacobi=True):
        """(x, iterations, final recurrence ‖r‖, ‖b‖); max_it < 0 runs exactly −max_it iterations"""
        nx, ny = b.shape
        hx, hy, bb = _f64(hx), _f64(hy), _f64(b)
        kp = _d(_f64(kappa)) if kappa is not None else None
        x = np.zeros((nx, ny))
        res = np.zeros(2)
        it = self.L.orc_fast_poisson2d_cg(ctypes.c_long(nx), ctypes.c_long(ny), _d(hx), _d(hy), kp, _d(bb),
                                          ctypes.c_double(tol), ctypes.c_int(max_it), ctypes.c_int(int(jacobi)),
                                          _d(x), _d(res))
        return x, it, res[0], res[1]

    def ms3d_cg(self, coords, mu, penalty, b, tol, max_it, jacobi=True):
        nx, ny, nz = (len(c) for c in coords)
        cs = [_f64(c) for c in coords]
        m, bb = _f64(mu), _f64(b)
        x = np.zeros(3 * nx * ny * nz)
        res = np.zeros(2)
        it = self.L.orc_fast_ms3d_cg(ctypes.c_long(nx), ctypes.c_long(ny), ctypes.c_long(nz), _d(cs[0]), _d(cs[1]),
                                     _d(cs[2]), _d(m), ctypes.c_double(penalty), _d(bb), ctypes.c_double(tol),
                                     ctypes.c_int(max_it), ctypes.c_int(int(jacobi)), _d(x), _d(res))
        return x, it, res[0], res[1]


def build_native(directory):
    """Builds the CPU baselines with -O3 -march=native into `directory` on
    this host; returns the library path, or None when the build fails."""
    out = Path(directory)
    try:
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "native", f"NATIVE_DIR={out}/"], check=True,
                       capture_output=True, timeout=300)
    except (subprocess.CalledProcessError, subprocess.TimeoutExpired, OSError):
        return None
    path = out / "libcpubaseline.so"
    return path if path.exists() else None


def threads():
    return int(lib().orc_threads())


CRITERIA = {"potential": 0, "moments": 1, "potential_and_moments": 2}


class Magnetostatics2D:
    """the 2D A_z operator (linear when ferro is None) on an nx × ny node
    grid, rows i + nx·j"""

    def __init__(self, shape, mu, ferro=None, bh_b=None, bh_h=None, criterion="moments"):
        self.nx, self.ny = shape
        self.mu = _f64(mu)
        self.ferro = _f64(ferro) if ferro is not None else None
        self.bh_b = _f64(bh_b if bh_b is not None else [0.0, 1.0])
        self.bh_h = _f64(bh_h if bh_h is not None else [0.0, 1.0])
        self.criterion = CRITERIA[criterion]
        self.n = self.nx * self.ny

    def _head(self):
        return [ctypes.c_long(self.nx), ctypes.c_long(self.ny), _d(self.mu),
                _d(self.ferro) if self.ferro is not None else None, _d(self.bh_b), _d(self.bh_h),
                ctypes.c_int(len(self.bh_b)), ctypes.c_int(self.criterion)]

    def apply(self, x, state=None):
        out = np.zeros(self.n)
        xx = _f64(x)
        st = _f64(state) if state is not None else None
        lib().orc_ms2d_apply(*self._head(), _d(st) if st is not None else None, _d(xx), _d(out))
        return out

    def diagonal(self, state=None):
        out = np.zeros(self.n)
        st = _f64(state) if state is not None else None
        lib().orc_ms2d_diagonal(*self._head(), _d(st) if st is not None else None, _d(out))
        return out

    def csr(self, state=None):
        """(row_ptr int64, cols int32, vals float64) of the operator (linear)
        or of its Jacobian at `state`, assembled as the reference does"""
        rp = np.zeros(self.n + 1, dtype=np.int64)
        cols = np.zeros(8 * self.n, dtype=np.int32)
        vals = np.zeros(8 * self.n)
        st = _f64(state) if state is not None else None
        lib().orc_ms2d_assemble_csr.restype = ctypes.c_long
        nnz = lib().orc_ms2d_assemble_csr(*self._head(), _d(st) if st is not None else None, rp.ctypes.data_as(_lp),
                                          cols.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _d(vals))
        return rp, cols[:nnz].copy(), vals[:nnz].copy()

    def newton(self, b, tol, max_inner, jacobi=True):
        x = np.zeros(self.n)
        out = np.zeros(4)
        bb = _f64(b)
        pc = 2 if jacobi == "spd_isai" else (1 if jacobi else 0)
        ok = lib().orc_ms2d_newton(*self._head(), _d(bb), ctypes.c_double(tol), ctypes.c_int(max_inner),
                                   ctypes.c_int(pc), _d(x), _d(out))
        return x, {"converged": bool(ok), "newton_steps": int(out[0]), "iterations": int(out[1]),
                   "initial_residual_l2": out[2], "final_residual_l2": out[3]}


def ms2d_rhs(coords, cell_jz, criterion="moments"):
    xc, yc = (_f64(c) for c in coords)
    out = np.zeros(len(xc) * len(yc))
    jz = _f64(cell_jz)
    lib().orc_ms2d_rhs(ctypes.c_long(len(xc)), ctypes.c_long(len(yc)), _d(xc), _d(yc), _d(jz),
                       ctypes.c_int(CRITERIA[criterion]), _d(out))
    return out


def ms2d_nodal_mu(shape, cell_mu, mu0):
    out = np.zeros(shape[0] * shape[1])
    cm = _f64(cell_mu)
    lib().orc_ms2d_nodal_mu(ctypes.c_long(shape[0]), ctypes.c_long(shape[1]), _d(cm), ctypes.c_double(mu0), _d(out))
    return out


def ms2d_post_process(coords, potential, cell_mu, cell_nonlinear=None, bh_b=None, bh_h=None):
    """cell induction, field, Maxwell stress and force density of a 2D A_z
    potential (cells x fastest)"""
    xc, yc = (_f64(c) for c in coords)
    ncell = (len(xc) - 1) * (len(yc) - 1)
    pot, mu = _f64(potential), _f64(cell_mu)
    nl = _f64(cell_nonlinear) if cell_nonlinear is not None else None
    bb = _f64(bh_b if bh_b is not None else [0.0])
    hh = _f64(bh_h if bh_h is not None else [0.0])
    out = [np.zeros((ncell, w)) for w in (3, 3, 6, 3)]
    lib().orc_ms2d_post_process(ctypes.c_long(len(xc)), ctypes.c_long(len(yc)), _d(xc), _d(yc), _d(pot), _d(mu),
                                _d(nl) if nl is not None else None, _d(bb), _d(hh),
                                ctypes.c_int(len(bb) if bh_b is not None else 0), *[_d(o) for o in out])
    return tuple(out)
