"""Python view of a batch of independent FSLBM domains (C ABI
fslbm_batch_*): one persistent launch per step over all admitted
domains, HBM-budget admission, optional hipGraph replay."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check_batch, lib
from .solver import RunParams


class BudgetFull(Exception):
    """Admission refused: the batch's HBM budget cannot hold the domain."""


class Batch:
    def __init__(self, precision: str = "f32", d1: int = 2, hbm_budget: int = 64 << 30):
        self._h = C.c_void_p()
        self.precision = precision
        check_batch(lib().fslbm_batch_create(_lib.PRECISIONS[precision], int(d1),
                                             int(hbm_budget), C.byref(self._h)))
        self.shapes = {}

    def close(self):
        if self._h:
            lib().fslbm_batch_destroy(self._h)
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

    def _chk(self, rc):
        return check_batch(rc, self._h)

    @staticmethod
    def domain_bytes(precision: str, shape, masked: bool) -> int:
        nz, ny, nx = shape
        m = np.zeros(1, dtype=np.uint8) if masked else None
        return lib().fslbm_batch_domain_bytes(_lib.PRECISIONS[precision],
                                              None if m is None else m.ctypes.data_as(C.c_void_p),
                                              nx, ny, nz)

    def add(self, params: RunParams, state, mask=None) -> int:
        nz, ny, nx = np.shape(state[0])
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in state]
        cp = params.to_c()
        if mask is not None:
            mask = np.ascontiguousarray(mask, dtype=np.uint8)
            mptr = mask.ctypes.data_as(C.c_void_p)
        else:
            mptr = None
        rc = lib().fslbm_batch_add(self._h, C.byref(cp), mptr, nx, ny, nz, *arrs)
        if rc == _lib.EBUSY:
            raise BudgetFull(lib().fslbm_batch_last_error(self._h).decode())
        slot = self._chk(rc)
        self.shapes[slot] = (nz, ny, nx)
        return slot

    def remove(self, slot: int) -> None:
        self._chk(lib().fslbm_batch_remove(self._h, int(slot)))
        self.shapes.pop(slot, None)

    def step(self, nsteps: int = 1) -> None:
        self._chk(lib().fslbm_batch_step(self._h, int(nsteps)))

    def sync(self) -> None:
        self._chk(lib().fslbm_batch_sync(self._h))

    def get_state(self, slot: int):
        shape = self.shapes[slot]
        n = int(np.prod(shape))
        out = [np.empty(n) for _ in range(4)]
        self._chk(lib().fslbm_batch_get_state(self._h, int(slot), *out))
        return [a.reshape(shape) for a in out]

    def query(self):
        used, budget = C.c_size_t(), C.c_size_t()
        nd, cells = C.c_int(), C.c_longlong()
        self._chk(lib().fslbm_batch_query(self._h, C.byref(used), C.byref(budget),
                                          C.byref(nd), C.byref(cells)))
        return {"used": used.value, "budget": budget.value, "domains": nd.value,
                "cells": cells.value}

    def use_graph(self, enable: bool = True) -> None:
        self._chk(lib().fslbm_batch_use_graph(self._h, int(bool(enable))))

    def set_chunk(self, chunk: int) -> None:
        self._chk(lib().fslbm_batch_set_chunk(self._h, int(chunk)))

    def autotune(self, trial_steps: int = 3):
        best, ms = C.c_int(), C.c_double()
        self._chk(lib().fslbm_batch_autotune(self._h, int(trial_steps), C.byref(best),
                                             C.byref(ms)))
        return best.value, ms.value

    def time_steps(self, nsteps: int) -> float:
        ms = C.c_double()
        self._chk(lib().fslbm_batch_time_steps(self._h, int(nsteps), C.byref(ms)))
        return ms.value
