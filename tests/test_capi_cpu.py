"""CPU-side checks of the C ABI: the library loads, exports every entry
point include/sil_hip.h declares, and rejects invalid requests with a
status and a message before touching the GPU."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_entry_points():
    names = set()
    for header in (ROOT / "include").glob("*.h"):
        text = header.read_text()
        names |= set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(sil_\w+)\s*\(", text, re.M))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    from similie_amd import _lib as L

    return L.lib()


def test_every_declared_entry_point_is_exported(lib):
    names = declared_entry_points()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_the_header():
    from similie_amd import _lib as L

    assert set(declared_entry_points()) <= set(L.SIGNATURES)


def test_abi_version(lib):
    assert lib.sil_abi_version() == 1


def test_invalid_requests_fail_before_any_gpu_call(lib):
    from similie_amd import _lib as L

    v = L.View()
    v.ndim = 2
    v.extent[0] = v.extent[1] = 4
    fake = ctypes.c_void_p(0x1000)
    # degree 2 has no coboundary in 2 dimensions
    assert lib.sil_coboundary(2, fake, ctypes.byref(v), fake, ctypes.byref(v), None) == -1
    assert b"coboundary" in lib.sil_last_error()
    bad = L.View()
    bad.ndim = 7
    assert lib.sil_coboundary(0, fake, ctypes.byref(bad), fake, ctypes.byref(v), None) == -1
    h = ctypes.c_void_p()
    ext = (ctypes.c_int64 * 1)(4)
    assert lib.sil_grid_create(ctypes.byref(h), 0, ext, None, None, 0, None) == -1
    assert lib.sil_grid_create(ctypes.byref(h), 1, ext, None, None, 0, None) == -1  # no coordinates
    assert lib.sil_operator_size(None) == -1
    st = L.SolverSettings()
    lib.sil_default_solver_settings(ctypes.byref(st))
    assert st.max_iterations == 2000 and st.relative_tolerance == 1e-12
    assert st.preconditioner == L.PRECOND_JACOBI
    dg = L.SolverDiagnostics()
    assert lib.sil_minimize_strong_formulation_residual(None, None, None, ctypes.byref(st), ctypes.byref(dg), None) == -1


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from similie_amd import _lib as L

    monkeypatch.setattr(L, "LIB_PATH", tmp_path / "absent.so")
    monkeypatch.setattr(L, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        L.lib()
