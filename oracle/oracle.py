"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the reference that GPU results are
compared with.  The product library never loads it.

Arrays are numpy arrays of shape (nz, ny, nx) (x fastest); a state is a
list [p, ux, uy, uz].
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

# FSLBM_ORACLE_SO selects another build of the same sources (the
# sanitizer build of tests/test_oracle_sanitized.py).
_SO = Path(os.environ.get("FSLBM_ORACLE_SO",
                          Path(__file__).resolve().parent / "_build" / "liboracle.so"))
_lib = None


class OrcParams(C.Structure):
    _fields_ = [("tblend", C.c_double), ("pin", C.c_double), ("nu", C.c_double),
                ("blend", C.c_double), ("d1", C.c_int)]


def lib():
    global _lib
    if _lib is None:
        if not _SO.exists():
            raise RuntimeError(f"{_SO} missing: run `make -C oracle`")
        L = C.CDLL(str(_SO))
        d, i = C.c_double, C.c_int
        vp = C.c_void_p
        P = C.POINTER(OrcParams)
        for name, args in {
            "orc_step_padded_f64": [i, i, i, i, vp, P, d, vp, vp],
            "orc_step_padded_f32": [i, i, i, i, vp, P, d, vp, vp],
            "orc_step_folded_f64": [i, i, i, vp, P, d, vp, vp],
            "orc_step_folded_f32": [i, i, i, vp, P, d, vp, vp],
            "orc_monitor_padded_f64": [i, i, i, i, d, vp, vp],
            "orc_stats_accumulate": [i, i, vp, vp, C.POINTER(i), C.POINTER(i)],
            "orc_time_folded_f32": [i, i, i, vp, P, i, vp, vp, C.POINTER(i)],
            "orc_time_folded_f64": [i, i, i, vp, P, i, vp, vp, C.POINTER(i)],
        }.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = d if name.startswith("orc_time") else None
        L.orc_max_threads.restype = i
        _lib = L
    return _lib


def params(tblend=1.0, pin=0.0, nu=0.01, blend=0.0, d1=2) -> OrcParams:
    return OrcParams(float(tblend), float(pin), float(nu), float(blend), int(d1))


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _pack(state, dtype):
    return np.ascontiguousarray(np.stack([np.asarray(s, dtype=dtype) for s in state]))


def step_folded(state, prm: OrcParams, mask=None, gauge=0.0, dtype=np.float64,
                nsteps: int = 1):
    """nsteps folded steps; the gauge applies to the first step only."""
    nz, ny, nx = np.shape(state[0])
    a = _pack(state, dtype)
    b = np.empty_like(a)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    fn = lib().orc_step_folded_f64 if dtype == np.float64 else lib().orc_step_folded_f32
    for s in range(nsteps):
        fn(nx, ny, nz, None if m is None else _ptr(m), C.byref(prm),
           float(gauge if s == 0 else 0.0), _ptr(a), _ptr(b))
        a, b = b, a
    return [a[f] for f in range(4)]


def step_padded(state, prm: OrcParams, IO, d2: int, p2=0.0, dtype=np.float64):
    """One reference-layout step (FSLBMcalc + bcond with gauge p2)."""
    lz, ly, lx = np.shape(state[0])
    a = _pack(state, dtype)
    b = np.empty_like(a)
    m = np.ascontiguousarray(IO, dtype=np.uint8)
    fn = lib().orc_step_padded_f64 if dtype == np.float64 else lib().orc_step_padded_f32
    fn(lx, ly, lz, d2, _ptr(m), C.byref(prm), float(p2), _ptr(a), _ptr(b))
    return [b[f] for f in range(4)]


def monitor_padded(state, d2: int, l1: float):
    lz, ly, lx = np.shape(state[0])
    a = _pack(state, np.float64)
    out = np.zeros(4)
    lib().orc_monitor_padded_f64(lx, ly, lz, d2, float(l1), _ptr(a), _ptr(out))
    return out


class Stats:
    """Reference statistics accumulator (FSLBM.cu:683-764)."""

    def __init__(self, n: int):
        self.n = n
        self.acc = np.zeros(11 * n)
        self.cnt = C.c_int(0)
        self.cnt2 = C.c_int(0)

    def accumulate(self, sample4, delta: int):
        s = np.ascontiguousarray(np.stack([np.asarray(x, dtype=np.float64).reshape(-1)
                                           for x in sample4]))
        lib().orc_stats_accumulate(self.n, int(delta), _ptr(self.acc), _ptr(s),
                                   C.byref(self.cnt), C.byref(self.cnt2))

    def read(self):
        return self.acc.reshape(11, self.n).copy(), self.cnt.value, self.cnt2.value


def time_folded(state, prm: OrcParams, mask=None, nsteps: int = 1, dtype=np.float32):
    """Seconds for nsteps folded steps and the OpenMP thread count."""
    nz, ny, nx = np.shape(state[0])
    a = _pack(state, dtype)
    b = np.empty_like(a)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    thr = C.c_int()
    fn = lib().orc_time_folded_f32 if dtype == np.float32 else lib().orc_time_folded_f64
    sec = fn(nx, ny, nz, None if m is None else _ptr(m), C.byref(prm), int(nsteps),
             _ptr(a), _ptr(b), C.byref(thr))
    return sec, thr.value


def max_threads() -> int:
    return lib().orc_max_threads()
