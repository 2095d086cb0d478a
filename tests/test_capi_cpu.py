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


def form_view(extents, ncomp):
    from similie_amd import _lib as L

    v = L.View()
    v.ndim = len(extents)
    s = ncomp
    for a in range(len(extents) - 1, -1, -1):
        v.extent[a] = extents[a]
        v.stride[a] = s
        s *= extents[a]
    v.comp_stride = 1
    v.comp_extent = ncomp
    return v


def test_component_extents_are_checked(lib):
    """A k-form must hold C(N, k) components along its component axis, and
    the output the count of the operator's output degree: a wrongly shaped
    array is refused before any kernel could index past its end."""
    fake = ctypes.c_void_p(0x1000)
    ok_in, ok_out = form_view([5, 6, 7], 3), form_view([5, 6, 7], 3)
    short_in, short_out = form_view([5, 6, 7], 1), form_view([5, 6, 7], 2)
    # 1-form → 2-form in 3D: 3 components each
    for fn in (lib.sil_coboundary, lib.sil_transposed_coboundary):
        assert fn(1, fake, ctypes.byref(short_in), fake, ctypes.byref(ok_out), None) == -1
        assert b"input holds 1 components" in lib.sil_last_error()
        assert fn(1, fake, ctypes.byref(ok_in), fake, ctypes.byref(short_out), None) == -1
        assert b"output holds 2 components" in lib.sil_last_error()
    # 0-form → 1-form: 1 and 3 components
    assert lib.sil_coboundary(0, fake, ctypes.byref(form_view([5, 6, 7], 1)), fake, ctypes.byref(short_in),
                              None) == -1
    assert b"a 1-form in 3 dimensions has 3" in lib.sil_last_error()


def test_python_views_carry_the_component_extent():
    import torch

    from similie_amd import exterior as E

    t = torch.zeros((4, 5, 6, 3), dtype=torch.float64)
    assert E.make_view(t).comp_extent == 3
    assert E.make_view(t.permute(3, 0, 1, 2), comp_dim=0).comp_extent == 3
    assert E.make_view(torch.zeros((4, 5), dtype=torch.float64), comp_dim=None).comp_extent == 1


@pytest.mark.parametrize("env,expect", [
    ({}, (1.0, False, 1.0)),
    ({"SIMILIE_MAGNETIC_Y_RESPONSE_SIGN_3D": "-1", "SIMILIE_DIVERGENCE_GAUGE_3D": "1",
      "SIMILIE_VECTOR_POTENTIAL_GAUGE_SIGN_3D": "-0.5"}, (-1.0, True, -1.0)),
    ({"SIMILIE_MAGNETIC_Y_RESPONSE_SIGN_3D": "abc", "SIMILIE_DIVERGENCE_GAUGE_3D": "0",
      "SIMILIE_VECTOR_POTENTIAL_GAUGE_SIGN_3D": "0"}, (1.0, False, 1.0)),
    ({"SIMILIE_MAGNETIC_Y_RESPONSE_SIGN_3D": "", "SIMILIE_DIVERGENCE_GAUGE_3D": "yes",
      "SIMILIE_VECTOR_POTENTIAL_GAUGE_SIGN_3D": "2.5"}, (1.0, True, 1.0)),
    ({"SIMILIE_DIVERGENCE_GAUGE_3D": ""}, (1.0, False, 1.0)),
])
def test_switches_from_the_environment(lib, monkeypatch, env, expect):
    """the reference's parsing (magnetostatics_onelab.hpp:796-828): a sign is
    +1 unless the variable holds a number below zero; the gauge is on when set,
    non-empty and not starting with '0'"""
    from similie_amd import solvers as S

    for name in ("SIMILIE_MAGNETIC_Y_RESPONSE_SIGN_3D", "SIMILIE_DIVERGENCE_GAUGE_3D",
                 "SIMILIE_VECTOR_POTENTIAL_GAUGE_SIGN_3D"):
        monkeypatch.delenv(name, raising=False)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    sw = S.ms3d_switches_from_environment()
    assert (sw["y_response_sign"], sw["divergence_gauge"], sw["gauge_sign"]) == expect


def test_switches_reject_unknown_criterion(lib):
    from similie_amd import _lib as L

    sw = L.Ms3dSwitches(1.0, 0, 1.0, 7)
    assert lib.sil_magnetostatics3d_set_switches(None, ctypes.byref(sw)) == L.SIL_ERR_INVALID
