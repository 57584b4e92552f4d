"""Python view of one FSLBM context (libfslbm.so through ctypes).

``Solver`` mirrors the reference's per-step hot path (FSLBM.cu:622-628:
FSLBMcalc slabs, shiftback, bcond) as ``step()``, its monitoring
(FSLBM.cu:630-666) as ``monitor()`` and its plane statistics
(FSLBM.cu:683-764) as ``stats_*``.  All device work runs in the HIP
kernels of the library; this module only moves host arrays across the
C ABI.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import Params, check, lib


@dataclass
class RunParams:
    """The six values of param.txt (pipe.m:67-69, FSLBM.cu:502-507)."""

    tblend: float = 1.0
    pin: float = 0.003
    nu: float = 40.06e-4
    d1: int = 2
    delta: int = 1
    blend: float = 0.0

    def to_c(self) -> Params:
        return Params(float(self.tblend), float(self.pin), float(self.nu),
                      float(self.blend), int(self.d1), int(self.delta))


class Solver:
    """One device-resident FSLBM domain in the folded layout.

    mask: uint8 array of shape (nz, ny, nx) (x fastest in memory) with codes
    0 = fluid and 125 = wall, or None for a fully periodic box.
    precision: "f64", "f32", "f16s" or "bf16s".
    """

    def __init__(self, shape, params: RunParams, mask=None, precision="f32",
                 kernel: int = _lib.KERNEL_AUTO):
        nz, ny, nx = shape
        self.shape = (nz, ny, nx)
        self.n = nx * ny * nz
        self.precision = precision
        L = lib()
        self._h = C.c_void_p()
        cp = params.to_c()
        if mask is not None:
            mask = np.ascontiguousarray(mask, dtype=np.uint8)
            if mask.size != self.n:
                raise ValueError("mask size does not match the box")
            mptr = mask.ctypes.data_as(C.c_void_p)
        else:
            mptr = None
        check(L.fslbm_create(C.byref(cp), mptr, nx, ny, nz,
                             _lib.PRECISIONS[precision], C.byref(self._h)))
        self.params = params
        self.mask = mask
        if kernel != _lib.KERNEL_AUTO:
            self.set_kernel(kernel)

    # -- lifetime -----------------------------------------------------------
    def close(self) -> None:
        if self._h:
            lib().fslbm_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc: int) -> int:
        return check(rc, self._h)

    # -- configuration --------------------------------------------------------
    def set_params(self, params: RunParams) -> None:
        cp = params.to_c()
        self._chk(lib().fslbm_set_params(self._h, C.byref(cp)))
        self.params = params

    def set_kernel(self, variant: int) -> None:
        self._chk(lib().fslbm_set_kernel(self._h, int(variant)))

    @property
    def kernel(self) -> int:
        return self._chk(lib().fslbm_get_kernel(self._h))

    def set_tile_chunk(self, chunk: int) -> None:
        self._chk(lib().fslbm_set_tile_chunk(self._h, int(chunk)))

    @property
    def tile_chunk(self) -> int:
        return self._chk(lib().fslbm_get_tile_chunk(self._h))

    @property
    def step_cells(self) -> int:
        """Lattice cells one step reads and writes (the tiled kernel skips
        tiles of wall cells)."""
        return self._chk(lib().fslbm_step_cells(self._h))

    def autotune(self, trial_steps: int = 5):
        """Pick the fastest tiled schedule; returns (chunk, ms per step).
        The state is not advanced."""
        best, ms = C.c_int(), C.c_double()
        self._chk(lib().fslbm_autotune(self._h, int(trial_steps), C.byref(best), C.byref(ms)))
        return best.value, ms.value

    # -- state ----------------------------------------------------------------
    def set_state(self, p, ux, uy, uz) -> None:
        arrs = [np.ascontiguousarray(a).reshape(-1) for a in (p, ux, uy, uz)]
        if arrs[0].dtype == np.float32:
            arrs = [a.astype(np.float32, copy=False) for a in arrs]
            self._chk(lib().fslbm_set_state_f32(self._h, *arrs))
        else:
            arrs = [a.astype(np.float64, copy=False) for a in arrs]
            self._chk(lib().fslbm_set_state(self._h, *arrs))

    def get_state(self, dtype=np.float64):
        out = [np.empty(self.n, dtype=dtype) for _ in range(4)]
        if dtype == np.float32:
            self._chk(lib().fslbm_get_state_f32(self._h, *out))
        else:
            self._chk(lib().fslbm_get_state(self._h, *out))
        return [a.reshape(self.shape) for a in out]

    def snapshot(self) -> bytes:
        """Exact snapshot of the stepping state (stored state, 16-bit
        storage scales, stochastic-rounding step counter, pending gauge)."""
        n = self._chk(lib().fslbm_snapshot_size(self._h))
        buf = C.create_string_buffer(n)
        self._chk(lib().fslbm_snapshot_save(self._h, buf, n))
        return buf.raw

    def load_snapshot(self, snap: bytes) -> None:
        """Restore a snapshot of a context with the same precision and
        cells: stepping then continues bitwise as from where it was taken."""
        buf = C.create_string_buffer(bytes(snap), len(snap))
        self._chk(lib().fslbm_snapshot_load(self._h, buf, len(snap)))

    # -- stepping ---------------------------------------------------------------
    def step(self, nsteps: int = 1) -> None:
        self._chk(lib().fslbm_step(self._h, int(nsteps)))

    def sync(self) -> None:
        self._chk(lib().fslbm_sync(self._h))

    def time_steps(self, nsteps: int) -> float:
        ms = C.c_double()
        self._chk(lib().fslbm_time_steps(self._h, int(nsteps), C.byref(ms)))
        return ms.value

    def set_gauge(self, p2: float) -> None:
        self._chk(lib().fslbm_set_gauge(self._h, float(p2)))

    # -- diagnostics --------------------------------------------------------------
    def monitor(self, set_gauge: bool = True):
        out = np.zeros(4)
        self._chk(lib().fslbm_monitor(self._h, int(bool(set_gauge)), out))
        return out

    def reduce(self):
        out = np.zeros(8)
        self._chk(lib().fslbm_reduce(self._h, out))
        return out

    def probe(self, x: int, y: int, z: int):
        out = np.zeros(4)
        self._chk(lib().fslbm_probe(self._h, x, y, z, out))
        return out

    def plane(self, axis: int, idx: int, d2: int, field: int):
        nz, ny, nx = self.shape
        n = ny * (nz + 2 * d2) if axis == 0 else nx * ny
        out = np.zeros(n)
        self._chk(lib().fslbm_get_plane(self._h, axis, idx, d2, field, out))
        return out

    def device_bytes(self) -> int:
        b = C.c_size_t()
        self._chk(lib().fslbm_device_bytes(self._h, C.byref(b)))
        return b.value

    # -- reference statistics ------------------------------------------------------
    def stats_config(self, xplane: int, d2: int) -> None:
        self._chk(lib().fslbm_stats_config(self._h, xplane, d2))

    def stats_accumulate(self, delta: int) -> None:
        self._chk(lib().fslbm_stats_accumulate(self._h, int(delta)))

    def stats_read(self):
        n = self._chk(lib().fslbm_stats_points(self._h))
        acc = np.zeros(11 * n)
        cnt, cnt2 = C.c_int(), C.c_int()
        self._chk(lib().fslbm_stats_read(self._h, acc.ctypes.data_as(C.c_void_p),
                                         C.byref(cnt), C.byref(cnt2)))
        return acc.reshape(11, n), cnt.value, cnt2.value

    def stats_load(self, acc, cnt: int, cnt2: int) -> None:
        a = np.ascontiguousarray(acc, dtype=np.float64).reshape(-1)
        self._chk(lib().fslbm_stats_load(self._h, a, int(cnt), int(cnt2)))

    # -- moments ---------------------------------------------------------------------
    def moments_reset(self) -> None:
        self._chk(lib().fslbm_moments_reset(self._h))

    def moments_accumulate(self) -> None:
        self._chk(lib().fslbm_moments_accumulate(self._h))

    def moments_read(self):
        nz, ny, nx = self.shape
        out = np.zeros(11 * nx * ny)
        s = C.c_long()
        self._chk(lib().fslbm_moments_read(self._h, out.ctypes.data_as(C.c_void_p),
                                           C.byref(s)))
        return out.reshape(11, ny, nx), s.value


def tile_plan(shape, precision: str = "f32", mask=None, blend: float = 0.0) -> dict:
    """The tiled kernel's schedule for a domain of shape (nz, ny, nx),
    computed on the host: tile rows, tiles holding fluid, cells one step
    moves and the default chunk.  `mask` is the z-invariant mask (one
    (ny, nx) plane or the full volume), None for a periodic box; `blend`
    the corrector's order-4 weight the domain is created with (FP16
    storage plans 12-row tiles in a walled box at blend 0, else 8)."""
    nz, ny, nx = shape
    mptr = None
    if mask is not None:
        m = np.ascontiguousarray(np.asarray(mask, dtype=np.uint8).reshape(-1, ny, nx)[0])
        mptr = m.ctypes.data_as(C.c_void_p)
    rows, ntiles, chunk = C.c_int(), C.c_int(), C.c_int()
    cells = C.c_longlong()
    check(lib().fslbm_tile_plan(nx, ny, nz, _lib.PRECISIONS[precision], float(blend), mptr,
                                C.byref(rows), C.byref(ntiles), C.byref(cells),
                                C.byref(chunk)))
    return {"tile_rows": rows.value, "tiles": ntiles.value, "cells": cells.value,
            "chunk": chunk.value}
