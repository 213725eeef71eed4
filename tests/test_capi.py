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
