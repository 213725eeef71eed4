"""ctypes binding of libvardyn_hip.so (the C ABI declared in include/vardyn.h).

The library is loaded from the package directory.  There is no fallback: if
the library is missing or a call fails, ``VardynError`` is raised.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libvardyn_hip.so"

VD_OK = 0
VD_F32, VD_F64 = 0, 1
VD_EULER, VD_RK2, VD_RK3 = 0, 1, 2
VD_REDUCE_SCRATCH = 1024

SCHEMES = {"Euler": VD_EULER, "rk2": VD_RK2, "rk3": VD_RK3}


class VardynError(RuntimeError):
    """A call into libvardyn_hip failed (status code and message)."""


class QgParams(C.Structure):
    _fields_ = [
        ("ny", C.c_int), ("nx", C.c_int), ("precision", C.c_int), ("time_scheme", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dt", C.c_double),
        ("g", C.c_double), ("f0", C.c_double), ("c", C.c_double),
        ("has_kdiffus", C.c_int), ("kdiffus", C.c_double),
        ("mask", C.c_void_p), ("beta", C.c_void_p), ("fbathy", C.c_void_p), ("mdt", C.c_void_p),
        ("dst_dense", C.c_int), ("gemm_config", C.c_int),
        ("nbatch", C.c_int), ("c_batch", C.c_void_p),
        ("dst_split", C.c_int),
    ]


class ObsTable(C.Structure):
    _fields_ = [
        ("n_obs", C.c_int), ("nnz", C.c_int), ("n_cols", C.c_int),
        ("row_ptr", C.c_void_p), ("col", C.c_void_p), ("w", C.c_void_p),
        ("y", C.c_void_p), ("err", C.c_void_p),
        ("tcol", C.c_void_p), ("tptr", C.c_void_p), ("trow", C.c_void_p), ("tw", C.c_void_p),
        ("partial", C.c_void_p),
    ]


class SwParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nl", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dt", C.c_double),
        ("H", C.c_void_p), ("g_prime", C.c_void_p), ("f", C.c_void_p), ("mask", C.c_void_p),
        ("slip_coef", C.c_double), ("bottom_drag_coef", C.c_double),
        ("h_min", C.c_double), ("h_min_sharpness", C.c_double),
        ("visc_coef", C.c_double), ("diff_coef", C.c_double),
        ("wind_u", C.c_void_p), ("wind_v", C.c_void_p),
    ]


# name -> (restype, argtypes)
_VP, _I, _D, _SZ, _I64 = C.c_void_p, C.c_int, C.c_double, C.c_size_t, C.c_int64
SIGNATURES = {
    "vd_version": (_I, [C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "vd_last_error": (C.c_char_p, []),
    "vd_qg_plan_create": (_I, [C.POINTER(QgParams), C.POINTER(_VP)]),
    "vd_qg_plan_destroy": (_I, [_VP]),
    "vd_qg_tape_bytes": (_I, [_VP, _I, C.POINTER(_SZ)]),
    "vd_qg_plan_bytes": (_I, [_VP, C.POINTER(_SZ)]),
    "vd_qg_autotune": (_I, [_VP, _I, C.POINTER(C.c_int), _VP]),
    "vd_qg_debug_scratch": (_I, [_VP, _I, C.POINTER(_VP), C.POINTER(_I), C.POINTER(_I)]),
    "vd_qg_step": (_I, [_VP, _VP, _VP, _VP, _I, _VP, _VP]),
    "vd_qg_step_tgl": (_I, [_VP, _VP, _VP, _VP, _VP, _VP, _I, _VP]),
    "vd_qg_step_adj": (_I, [_VP, _VP, _VP, _VP, _VP, _I, _VP, _I, _VP]),
    "vd_qg_pv2h": (_I, [_VP, _VP, _VP, _VP, _VP, _VP]),
    "vd_qg_h2pv": (_I, [_VP, _VP, _VP, _VP, _VP]),
    "vd_qg_rhs": (_I, [_VP, _VP, _VP, _VP, C.c_double, _VP, _VP]),
    "vd_qg_rhs_adj": (_I, [_VP, _VP, _VP, _VP, C.c_double, _VP, _VP, _VP, _VP]),
    "vd_qg_fill_land": (_I, [_VP, _VP, _D, _VP]),
    "vd_qg_prep": (_I, [_VP, _VP, _VP, _VP, _VP, _VP]),
    "vd_qg_finish": (_I, [_VP, _VP, _VP, _VP]),
    "vd_qg_adj_init": (_I, [_VP, _VP, _VP, _VP, _VP]),
    "vd_qg_adj_final": (_I, [_VP, _VP, _VP, _VP, _VP, _VP]),
    "vd_group_create": (_I, [C.POINTER(_VP), _I, C.POINTER(_VP)]),
    "vd_group_destroy": (_I, [_VP]),
    "vd_group_offsets": (_I, [_VP, C.POINTER(_SZ)]),
    "vd_group_stage": (_I, [_VP, _I, _VP, _VP, _VP, _VP, _VP, _D, _D, _D, _VP, _VP, _VP]),
    "vd_group_inversion": (_I, [_VP, _I, _VP, _VP, _VP, _VP, _I, _VP]),
    "vd_group_stage_adj": (_I, [_VP, _I, _VP, _VP, _VP, _VP, _VP, _D, _D, _D, _VP, _VP, _VP,
                                _VP, _VP, _VP]),
    "vd_group_prep": (_I, [_VP, _I, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "vd_group_finish": (_I, [_VP, _I, _VP, _VP, _VP]),
    "vd_group_adj_init": (_I, [_VP, _I, _VP, _VP, _VP, _VP]),
    "vd_group_adj_final": (_I, [_VP, _I, _VP, _VP, _VP, _VP, _VP]),
    "vd_group_autotune": (_I, [_VP, _I, C.POINTER(_I), _VP]),
    "vd_group_set_tuning": (_I, [_VP, C.POINTER(_I)]),
    "vd_obs_misfit": (_I, [C.POINTER(ObsTable), _I, _VP, _VP, _VP, _VP, _D, _VP]),
    "vd_obs_adjoint": (_I, [C.POINTER(ObsTable), _I, _VP, _VP, _VP, _VP]),
    "vd_obs_grid_create": (_I, [_VP, C.c_int64, _D, C.POINTER(_VP)]),
    "vd_obs_grid_destroy": (_I, [_VP]),
    "vd_obs_grid_query": (_I, [_VP, _VP, _I, _I, _VP, _VP, _VP, _VP]),
    "vd_tiles_gather": (_I, [_I, _I, _I, _I, _I, _I, _VP, _VP, _VP, _VP, _VP, _VP]),
    "vd_tiles_blend": (_I, [_I, _I, _I, _I, _I, _I, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "vd_tiles_blend_adj": (_I, [_I, _I, _I, _I, _I, _I, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "vd_tiles_gather_adj": (_I, [_I, _I, _I, _I, _I, _I, _VP, _VP, _VP, _VP, _VP]),
    "vd_bm_create": (_I, [_VP, _VP]),
    "vd_bm_destroy": (_I, [_VP]),
    "vd_bm_info": (_I, [_VP, C.POINTER(C.c_longlong), C.POINTER(C.c_int)]),
    "vd_bm_workspace_size": (_I, [_VP, _I, C.POINTER(_SZ)]),
    "vd_bm_operg": (_I, [_VP, _I, _VP, _VP, _VP, _VP, _VP]),
    "vd_bm_operg_adj": (_I, [_VP, _I, _VP, _VP, _VP, _VP, _VP]),
    "vd_bm_uv_tables": (_I, [_VP, _VP, _VP, _VP, _VP]),
    "vd_bm_ssh2uv": (_I, [_VP, _I, _VP, _VP, _VP, _VP]),
    "vd_bm_ssh2uv_adj": (_I, [_VP, _I, _VP, _VP, _VP, _VP]),
    "vd_w3_create": (_I, [_VP, _VP]),
    "vd_w3_destroy": (_I, [_VP]),
    "vd_w3_info": (_I, [_VP, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "vd_w3_workspace_size": (_I, [_VP, C.POINTER(_SZ)]),
    "vd_w3_operg": (_I, [_VP, _I, _VP, _VP, _VP, _VP, _VP]),
    "vd_w3_operg_adj": (_I, [_VP, _I, _VP, _VP, _VP, _VP, _VP]),
    "vd_sw_create": (_I, [C.POINTER(SwParams), C.POINTER(_VP)]),
    "vd_sw_destroy": (_I, [_VP]),
    "vd_sw_info": (_I, [_VP, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "vd_sw_rhs": (_I, [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "vd_sw_step": (_I, [_VP] * 10 + [_I, _VP]),
    "vd_sw_step_tgl": (_I, [_VP] * 16 + [_I, _VP]),
    "vd_sw_step_adj": (_I, [_VP] * 16 + [_I, _VP]),
    "vd_sw_step_keep": (_I, [_VP] * 10 + [_I, _I, _VP, _VP]),
    "vd_sw_step_adj_kept": (_I, [_VP] * 5 + [_I] + [_VP] * 9 + [_I, _VP]),
    "vd_merge_create": (_I, [_VP, _VP]),
    "vd_merge_destroy": (_I, [_VP]),
    "vd_merge_info": (_I, [_VP, C.POINTER(C.c_longlong)]),
    "vd_merge_weights": (_I, [_VP, _VP, _VP, _VP]),
    "vd_merge_apply": (_I, [_VP, _I, _VP, _VP, _VP]),
    "vd_merge_blend": (_I, [C.c_longlong, C.c_double, C.c_double, _VP, _VP, _VP, _VP]),
    "vd_dot_f64": (_I, [_I64, _VP, _VP, _D, _VP, _VP, _VP]),
    "vd_dot_f64_noalloc": (_I, [_I64, _VP, _VP, _D, _VP, _VP]),
    "vd_axpby": (_I, [_I, _I64, _D, _VP, _D, _VP, _VP]),
    "vd_convert": (_I, [_I, _VP, _I, _VP, _I64, _I, _VP]),
}

_lib = None


def load():
    """Load libvardyn_hip.so once; raise VardynError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise VardynError(f"{LIB_PATH} is not built (run `python -m vardyn_amd.build`)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    return load().vd_last_error().decode()


def check(rc: int, what: str = "") -> None:
    if rc != VD_OK:
        raise VardynError(f"{what} failed with status {rc}: {last_error()}")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)
