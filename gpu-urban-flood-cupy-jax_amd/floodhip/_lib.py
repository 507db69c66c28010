"""ctypes binding of libfloodhip.so (C ABI declared in include/floodhip.h).

The shared library is loaded from this package directory only; there is no
fallback: a missing library raises BackendUnavailableError.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

from .errors import BackendUnavailableError

LIB_PATH = Path(__file__).resolve().parent / "libfloodhip.so"

FH_OK, FH_ERR_VALUE, FH_ERR_SIMULATION, FH_ERR_BUDGET, FH_ERR_UNAVAILABLE, FH_ERR_RUNTIME = range(6)
LEDGER_TERMS = 13


class FhParams(C.Structure):
    _fields_ = [
        ("gravity", C.c_double), ("min_depth", C.c_double), ("cfl", C.c_double),
        ("reuse_factor", C.c_double), ("max_dt", C.c_double), ("dt_floor", C.c_double),
        ("max_speed", C.c_double), ("inlet_orifice_q", C.c_double), ("slot_wave_speed", C.c_double),
        ("fixed_dt", C.c_double), ("cfl_interval", C.c_int32), ("compensated", C.c_int32),
        ("precision", C.c_int32), ("use_graph", C.c_int32), ("steps_per_launch", C.c_int32),
        ("block_size", C.c_int32), ("xcd_remap", C.c_int32), ("device", C.c_int32),
        ("edges_per_thread", C.c_int32), ("pack_order", C.c_int32),
        ("cell_order", C.c_int32), ("source_classes", C.c_int32),
        ("surface_mode", C.c_int32), ("share_mesh", C.c_int32),
    ]


class FhClock(C.Structure):
    _fields_ = [
        ("time_s", C.c_double), ("cached_dt_s", C.c_double), ("last_dt_s", C.c_double),
        ("global_step", C.c_int64), ("interval_steps", C.c_int64), ("cfl_evaluations", C.c_int64),
    ]


_f = C.POINTER(C.c_float)
_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int32)
_u8 = C.POINTER(C.c_uint8)
_i64 = C.c_int64
_ctx = C.c_void_p

SIGNATURES = {
    "fh_default_params": (None, [C.POINTER(FhParams)]),
    "fh_version": (C.c_char_p, []),
    "fh_device_count": (C.c_int, []),
    "fh_create": (_ctx, [C.POINTER(FhParams)]),
    "fh_create_batch": (_ctx, [C.POINTER(FhParams), C.c_int32]),
    "fh_select_member": (C.c_int, [_ctx, C.c_int32]),
    "fh_member_count": (C.c_int32, [_ctx]),
    "fh_destroy": (None, [_ctx]),
    "fh_last_error": (C.c_char_p, [_ctx]),
    "fh_global_error": (C.c_char_p, []),
    "fh_upload_surface": (C.c_int, [_ctx, _i64, _i64, _f, _f, _f, _f, _i, _i, _f, _f, _f, _i, _u8]),
    "fh_upload_coupling": (C.c_int, [_ctx, _i64, _i64, _i64, _i, _i, _i, _f, _f, _f, _f, _f, _i, _i, _f,
                                     C.c_int32]),
    "fh_upload_pipe": (C.c_int, [_ctx, _i64, _i, _i, _f, _f, _f, _f, _f, _f, _i, _f, _f, _f, _f, _f, _f]),
    "fh_upload_recharge": (C.c_int, [_ctx, _i64, _i, _f, _f, _f, _f]),
    "fh_set_forcing": (C.c_int, [_ctx, _i64, _f, _f, _f, _f, _f, _f, _f, _i64, _f, _i64, _i, _f]),
    "fh_set_boundary_stage": (C.c_int, [_ctx, _i64, _f]),
    "fh_set_state": (C.c_int, [_ctx, _f, _f, _f, _f, _f, _f, _f]),
    "fh_get_state": (C.c_int, [_ctx, _f, _f, _f, _f, _f, _f, _f]),
    "fh_get_state_f64": (C.c_int, [_ctx, _d, _d, _d]),
    "fh_get_drainage_f64": (C.c_int, [_ctx, _d, _d]),
    "fh_set_clock": (C.c_int, [_ctx, C.c_double, C.c_double, _i64]),
    "fh_get_clock": (C.c_int, [_ctx, C.POINTER(FhClock)]),
    "fh_set_ledger": (C.c_int, [_ctx, _d]),
    "fh_get_ledger": (C.c_int, [_ctx, _d]),
    "fh_advance_to": (C.c_int, [_ctx, C.c_double, C.c_double, _i64, C.POINTER(FhClock), _d]),
    "fh_step": (C.c_int, [_ctx, C.c_double, C.c_double, _f, _d]),
    "fh_stable_timestep": (C.c_int, [_ctx, _d]),
    "fh_stable_timestep_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, _i64, C.c_double,
                                             C.c_void_p, C.c_void_p]),
    "fh_profile_kernels": (C.c_int, [_ctx, C.c_int32, C.c_int32, _i, C.c_char_p, _d, _d,
                                     C.POINTER(C.c_int64)]),
    "fh_device_bytes": (C.c_int64, [_ctx]),
    "fh_dry_weather_step": (C.c_int, [_ctx, C.c_double, _f, _f, C.c_int32, _d]),
    "fh_copy_probe": (C.c_int, [C.c_int64, C.c_int32, C.c_int32, _d]),
    "fh_division_selftest": (C.c_int, [C.POINTER(C.c_int64)]),
    "fh_pow_selftest": (C.c_int, [C.POINTER(C.c_int64)]),
}

_LIB = None


def load(path: Path | None = None):
    """Load the HIP library once; raises BackendUnavailableError if it is missing."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise BackendUnavailableError(
            f"{p} is missing: build it with `python -m floodhip.build` (hipcc, gfx950)")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
    return lib


def exported_symbols():
    return list(SIGNATURES)
