"""ctypes binding of the C ABI in include/fslbm.h (libfslbm.so).

The library is loaded from the package directory; a missing library is an
error, never a fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libfslbm.so"

F64, F32, F16S, BF16S = 0, 1, 2, 3
PRECISIONS = {"f64": F64, "f32": F32, "f16s": F16S, "bf16s": BF16S}
KERNEL_AUTO, KERNEL_PLAIN, KERNEL_TILED = 0, 1, 2

# Every symbol include/fslbm.h declares (checked by the CPU tests).
EXPORTS = [
    "fslbm_abi_version", "fslbm_error_string", "fslbm_create", "fslbm_destroy",
    "fslbm_last_error", "fslbm_set_params", "fslbm_get_params",
    "fslbm_set_kernel", "fslbm_get_kernel", "fslbm_set_tile_chunk",
    "fslbm_get_tile_chunk", "fslbm_step_cells", "fslbm_tile_plan", "fslbm_autotune", "fslbm_set_state",
    "fslbm_get_state", "fslbm_set_state_f32", "fslbm_get_state_f32",
    "fslbm_snapshot_size", "fslbm_snapshot_save", "fslbm_snapshot_load",
    "fslbm_step", "fslbm_set_gauge", "fslbm_monitor", "fslbm_reduce",
    "fslbm_probe", "fslbm_get_plane", "fslbm_stats_config",
    "fslbm_stats_accumulate", "fslbm_stats_read", "fslbm_stats_load",
    "fslbm_stats_points", "fslbm_moments_reset", "fslbm_moments_accumulate",
    "fslbm_moments_read", "fslbm_sync", "fslbm_time_steps",
    "fslbm_device_bytes", "fslbm_pipe_m_mask", "fslbm_circle_mask",
    "fslbm_init_poiseuille", "fslbm_init_taylor_green", "fslbm_fold_mask",
    "fslbm_batch_create", "fslbm_batch_destroy", "fslbm_batch_last_error",
    "fslbm_batch_domain_bytes", "fslbm_batch_add", "fslbm_batch_remove",
    "fslbm_batch_step", "fslbm_batch_get_state", "fslbm_batch_query",
    "fslbm_batch_use_graph", "fslbm_batch_set_chunk", "fslbm_batch_autotune",
    "fslbm_batch_time_steps", "fslbm_batch_sync",
    "fslbm_server_create", "fslbm_server_set_policy", "fslbm_server_destroy",
    "fslbm_server_last_error",
    "fslbm_server_clock", "fslbm_server_submit", "fslbm_server_pump",
    "fslbm_server_next_done", "fslbm_server_fetch", "fslbm_server_drop",
    "fslbm_server_query", "fslbm_step_noalloc",
]
EBUSY = -6
SERVE_SHARE, SERVE_SHORTEST = 0, 1  # serving policies
NOALLOC_MAX_COLUMNS = 1 << 19  # FSLBM_NOALLOC_MAX_COLUMNS


class Params(C.Structure):
    """fslbm_params: the six values of param.txt (pipe.m:67-69)."""

    _fields_ = [
        ("tblend", C.c_double),
        ("pin", C.c_double),
        ("nu", C.c_double),
        ("blend", C.c_double),
        ("d1", C.c_int),
        ("delta", C.c_int),
    ]


class Request(C.Structure):
    """fslbm_request: one solve of the serving loop."""

    _fields_ = [
        ("params", Params),
        ("nx", C.c_int),
        ("ny", C.c_int),
        ("nz", C.c_int),
        ("precision", C.c_int),
        ("nsteps", C.c_longlong),
    ]


class Completion(C.Structure):
    """fslbm_completion: a finished (or failed) request."""

    _fields_ = [
        ("id", C.c_longlong),
        ("status", C.c_int),
        ("steps", C.c_longlong),
        ("t_submit", C.c_double),
        ("t_admit", C.c_double),
        ("t_done", C.c_double),
    ]


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")

_lib = None


def lib() -> C.CDLL:
    """Load libfslbm.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'`")
    L = C.CDLL(str(LIB_PATH))
    vp, ip, i, d = C.c_void_p, C.POINTER(C.c_int), C.c_int, C.c_double
    sig = {
        "fslbm_abi_version": (i, []),
        "fslbm_error_string": (C.c_char_p, [i]),
        "fslbm_create": (i, [C.POINTER(Params), C.c_void_p, i, i, i, i, C.POINTER(vp)]),
        "fslbm_destroy": (None, [vp]),
        "fslbm_last_error": (C.c_char_p, [vp]),
        "fslbm_set_params": (i, [vp, C.POINTER(Params)]),
        "fslbm_get_params": (i, [vp, C.POINTER(Params)]),
        "fslbm_set_kernel": (i, [vp, i]),
        "fslbm_get_kernel": (i, [vp]),
        "fslbm_set_tile_chunk": (i, [vp, i]),
        "fslbm_get_tile_chunk": (i, [vp]),
        "fslbm_step_cells": (C.c_longlong, [vp]),
        "fslbm_tile_plan": (i, [i, i, i, i, d, C.c_void_p, ip, ip, C.POINTER(C.c_longlong), ip]),
        "fslbm_autotune": (i, [vp, i, ip, C.POINTER(d)]),
        "fslbm_set_state": (i, [vp, _dp, _dp, _dp, _dp]),
        "fslbm_get_state": (i, [vp, _dp, _dp, _dp, _dp]),
        "fslbm_set_state_f32": (i, [vp, _fp, _fp, _fp, _fp]),
        "fslbm_get_state_f32": (i, [vp, _fp, _fp, _fp, _fp]),
        "fslbm_snapshot_size": (C.c_longlong, [vp]),
        "fslbm_snapshot_save": (i, [vp, vp, C.c_longlong]),
        "fslbm_snapshot_load": (i, [vp, vp, C.c_longlong]),
        "fslbm_step": (i, [vp, i]),
        "fslbm_set_gauge": (i, [vp, d]),
        "fslbm_monitor": (i, [vp, i, _dp]),
        "fslbm_reduce": (i, [vp, _dp]),
        "fslbm_probe": (i, [vp, i, i, i, _dp]),
        "fslbm_get_plane": (i, [vp, i, i, i, i, _dp]),
        "fslbm_stats_config": (i, [vp, i, i]),
        "fslbm_stats_accumulate": (i, [vp, i]),
        "fslbm_stats_read": (i, [vp, C.c_void_p, ip, ip]),
        "fslbm_stats_load": (i, [vp, _dp, i, i]),
        "fslbm_stats_points": (i, [vp]),
        "fslbm_moments_reset": (i, [vp]),
        "fslbm_moments_accumulate": (i, [vp]),
        "fslbm_moments_read": (i, [vp, C.c_void_p, C.POINTER(C.c_long)]),
        "fslbm_sync": (i, [vp]),
        "fslbm_time_steps": (i, [vp, i, C.POINTER(d)]),
        "fslbm_device_bytes": (i, [vp, C.POINTER(C.c_size_t)]),
        "fslbm_pipe_m_mask": (C.c_long, [i, i, i, _u8]),
        "fslbm_circle_mask": (C.c_long, [i, i, i, d, _u8]),
        "fslbm_init_poiseuille": (i, [i, i, i, _u8, d, d, d, d, C.c_uint64, _dp, _dp, _dp, _dp]),
        "fslbm_init_taylor_green": (i, [i, i, i, d, d, _dp, _dp, _dp, _dp]),
        "fslbm_fold_mask": (i, [i, i, i, i, _u8, _u8]),
        "fslbm_batch_create": (i, [i, i, C.c_size_t, C.POINTER(vp)]),
        "fslbm_batch_destroy": (None, [vp]),
        "fslbm_batch_last_error": (C.c_char_p, [vp]),
        "fslbm_batch_domain_bytes": (C.c_longlong, [i, C.c_void_p, i, i, i]),
        "fslbm_batch_add": (i, [vp, C.POINTER(Params), C.c_void_p, i, i, i, _dp, _dp, _dp, _dp]),
        "fslbm_batch_remove": (i, [vp, i]),
        "fslbm_batch_step": (i, [vp, i]),
        "fslbm_batch_get_state": (i, [vp, i, _dp, _dp, _dp, _dp]),
        "fslbm_batch_query": (i, [vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t), ip,
                                  C.POINTER(C.c_longlong)]),
        "fslbm_batch_use_graph": (i, [vp, i]),
        "fslbm_batch_set_chunk": (i, [vp, i]),
        "fslbm_batch_autotune": (i, [vp, i, ip, C.POINTER(d)]),
        "fslbm_batch_time_steps": (i, [vp, i, C.POINTER(d)]),
        "fslbm_batch_sync": (i, [vp]),
        "fslbm_server_create": (i, [i, C.c_size_t, i, C.POINTER(vp)]),
        "fslbm_server_set_policy": (i, [vp, i, C.c_longlong]),
        "fslbm_server_destroy": (None, [vp]),
        "fslbm_server_last_error": (C.c_char_p, [vp]),
        "fslbm_server_clock": (d, [vp]),
        "fslbm_server_submit": (C.c_longlong, [vp, C.POINTER(Request), C.c_void_p,
                                               _dp, _dp, _dp, _dp]),
        "fslbm_server_pump": (i, [vp]),
        "fslbm_server_next_done": (i, [vp, C.POINTER(Completion)]),
        "fslbm_server_fetch": (i, [vp, C.c_longlong, _dp, _dp, _dp, _dp]),
        "fslbm_server_drop": (i, [vp, C.c_longlong]),
        "fslbm_server_query": (i, [vp, ip, ip, ip, C.POINTER(C.c_size_t),
                                   C.POINTER(C.c_size_t)]),
        "fslbm_step_noalloc": (i, [C.POINTER(Params), vp, i, i, i, i, vp, vp, d, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class FslbmError(RuntimeError):
    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


def check(rc: int, ctx=None) -> int:
    if rc < 0:
        L = lib()
        msg = L.fslbm_last_error(ctx).decode() if ctx is not None else \
            L.fslbm_last_error(None).decode()
        raise FslbmError(f"{L.fslbm_error_string(rc).decode()}: {msg}", rc)
    return rc


def check_batch(rc: int, b=None) -> int:
    if rc < 0:
        L = lib()
        msg = L.fslbm_batch_last_error(b).decode()
        raise FslbmError(f"{L.fslbm_error_string(rc).decode()}: {msg}", rc)
    return rc


def check_server(rc: int, srv=None) -> int:
    if rc < 0:
        L = lib()
        msg = L.fslbm_server_last_error(srv).decode()
        raise FslbmError(f"{L.fslbm_error_string(rc).decode()}: {msg}", rc)
    return rc
