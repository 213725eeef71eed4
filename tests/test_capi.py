"""The C-ABI library loads and exports every entry point that include/vardyn.h declares
(no compute calls: these run without a GPU)."""
import re
from pathlib import Path

import pytest

from vardyn_amd import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "vardyn.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "vd_qg_step" in names and "vd_qg_step_adj" in names and "vd_obs_misfit" in names


def test_library_exports_all_declared_symbols():
    if not _lib.LIB_PATH.exists():
        pytest.fail("libvardyn_hip.so is not built (python -m vardyn_amd.build)")
    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), f"{name} declared in vardyn.h but not exported"


def test_binding_covers_header():
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_version_and_error_without_gpu():
    import ctypes as C
    lib = _lib.load()
    ma, mi = C.c_int(), C.c_int()
    assert lib.vd_version(C.byref(ma), C.byref(mi)) == 0
    assert (ma.value, mi.value) == (0, 1)
    # argument validation happens before any device call
    assert lib.vd_qg_plan_create(None, None) == -1
    assert b"null" in lib.vd_last_error()
    p = _lib.QgParams()
    p.ny, p.nx, p.precision, p.time_scheme = 4, 4, 0, 1
    h = C.c_void_p()
    assert lib.vd_qg_plan_create(C.byref(p), C.byref(h)) == -2


def test_error_paths_without_device():
    """Null/invalid arguments are rejected before any device work, with a message."""
    import ctypes as C
    lib = _lib.load()
    null = C.c_void_p()
    assert lib.vd_qg_step(None, null, null, null, 1, None, None) == -1
    assert lib.vd_qg_step_adj(None, null, null, null, null, 1, null, 0, None) == -1
    assert lib.vd_qg_step_tgl(None, null, null, null, null, None, 1, None) == -1
    assert lib.vd_qg_pv2h(None, null, null, null, null, None) == -1
    assert lib.vd_obs_misfit(None, 0, null, null, null, None, 1.0, None) == -1
    assert lib.vd_obs_adjoint(None, 0, null, null, null, None) == -1
    assert lib.vd_dot_f64(10, None, None, 1.0, None, None, None) == -1
    assert lib.vd_axpby(0, 10, 1.0, None, 1.0, None, None) == -1
    assert lib.vd_convert(0, None, 1, None, 10, 0, None) == -1
    assert lib.vd_last_error()  # a message is set
    # observation-operator construction: grid and query arguments
    import numpy as np
    xyz = np.zeros((4, 3))
    g = C.c_void_p()
    assert lib.vd_obs_grid_create(None, 4, 10.0, C.byref(g)) == -1
    assert lib.vd_obs_grid_create(xyz.ctypes.data, 0, 10.0, C.byref(g)) == -1
    assert lib.vd_obs_grid_create(xyz.ctypes.data, 4, float("nan"), C.byref(g)) == -1
    assert b"dmax" in lib.vd_last_error()
    assert lib.vd_obs_grid_create(xyz.ctypes.data, 4, -1.0, C.byref(g)) == -1
    bad = np.array([[0.0, 0.0, np.inf]] * 4)
    assert lib.vd_obs_grid_create(bad.ctypes.data, 4, 10.0, C.byref(g)) == -1
    assert b"finite" in lib.vd_last_error()
    assert lib.vd_obs_grid_query(None, null, 1, 4, null, null, None, None) == -1
    # parameter validation of plan creation
    p = _lib.QgParams()
    p.ny, p.nx, p.precision, p.time_scheme = 64, 64, 7, 1
    p.dx = p.dy = 1e4
    p.dt, p.g, p.f0, p.c = 900.0, 9.81, 1e-4, 2.0
    h = C.c_void_p()
    assert lib.vd_qg_plan_create(C.byref(p), C.byref(h)) == -1
    p.precision, p.time_scheme = 1, 9
    assert lib.vd_qg_plan_create(C.byref(p), C.byref(h)) == -1
    p.time_scheme, p.c = 1, -1.0
    assert lib.vd_qg_plan_create(C.byref(p), C.byref(h)) == -1
    p.c, p.gemm_config = 2.0, 12
    assert lib.vd_qg_plan_create(C.byref(p), C.byref(h)) == -1
    assert b"gemm_config" in lib.vd_last_error()
    p.gemm_config, p.nbatch = 0, -2
    assert lib.vd_qg_plan_create(C.byref(p), C.byref(h)) == -1
    # split-product GEMMs: 5..7 digits, FP64 plans only, their own tile range
    p.nbatch, p.dst_split = 1, 4
    assert lib.vd_qg_plan_create(C.byref(p), C.byref(h)) == -1
    assert b"dst_split" in lib.vd_last_error()
    p.dst_split, p.precision = 6, 0
    assert lib.vd_qg_plan_create(C.byref(p), C.byref(h)) == -1
    assert b"FP64" in lib.vd_last_error()
    p.precision, p.gemm_config = 1, 5
    assert lib.vd_qg_plan_create(C.byref(p), C.byref(h)) == -1
    assert b"0..4" in lib.vd_last_error()
