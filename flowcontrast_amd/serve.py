"""Python view of the serving loop (C ABI fslbm_server_*): requests are
queued as they arrive, admitted in arrival order while the HBM budget
holds them, advanced together in batched launches and handed back when
their step budget is spent."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check_server, lib
from .solver import RunParams


@dataclass
class Completion:
    id: int
    status: int
    steps: int
    t_submit: float
    t_admit: float
    t_done: float

    @property
    def ok(self) -> bool:
        return self.status == 0

    @property
    def latency(self) -> float:
        """Time to solution: submission to hand-back, in seconds."""
        return self.t_done - self.t_submit


class Server:
    def __init__(self, d1: int = 2, hbm_budget: int = 64 << 30, quantum: int = 200):
        self._h = C.c_void_p()
        check_server(lib().fslbm_server_create(int(d1), int(hbm_budget), int(quantum),
                                               C.byref(self._h)))
        self.shapes = {}

    def close(self):
        if self._h:
            lib().fslbm_server_destroy(self._h)
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
        return check_server(rc, self._h)

    def set_policy(self, policy: str = "share", fill_cells: int = 0) -> None:
        """Which resident requests step in a pump: "share" (all of them, the
        default) or "shortest" (least work left first, until fill_cells
        cells; 0 = the library's default).  Results do not depend on the
        policy."""
        code = {"share": _lib.SERVE_SHARE, "shortest": _lib.SERVE_SHORTEST}[policy]
        check_server(lib().fslbm_server_set_policy(self._h, code, int(fill_cells)), self._h)

    def clock(self) -> float:
        return lib().fslbm_server_clock(self._h)

    def submit(self, params: RunParams, state, mask=None, precision: str = "f32",
               nsteps: int = 1000) -> int:
        nz, ny, nx = np.shape(state[0])
        req = _lib.Request(params.to_c(), nx, ny, nz, _lib.PRECISIONS[precision], int(nsteps))
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in state]
        if mask is not None:
            mask = np.ascontiguousarray(mask, dtype=np.uint8)
            mptr = mask.ctypes.data_as(C.c_void_p)
        else:
            mptr = None
        rid = self._chk(lib().fslbm_server_submit(self._h, C.byref(req), mptr, *arrs))
        self.shapes[rid] = (nz, ny, nx)
        return rid

    def pump(self) -> int:
        return self._chk(lib().fslbm_server_pump(self._h))

    def completions(self):
        """Pop every unreported completion."""
        out = []
        c = _lib.Completion()
        while self._chk(lib().fslbm_server_next_done(self._h, C.byref(c))) == 1:
            out.append(Completion(c.id, c.status, c.steps, c.t_submit, c.t_admit, c.t_done))
        return out

    def fetch(self, rid: int):
        shape = self.shapes[rid]
        n = int(np.prod(shape))
        out = [np.empty(n) for _ in range(4)]
        self._chk(lib().fslbm_server_fetch(self._h, int(rid), *out))
        return [a.reshape(shape) for a in out]

    def drop(self, rid: int) -> None:
        self._chk(lib().fslbm_server_drop(self._h, int(rid)))
        self.shapes.pop(rid, None)

    def query(self):
        q, r, f = C.c_int(), C.c_int(), C.c_int()
        used, budget = C.c_size_t(), C.c_size_t()
        self._chk(lib().fslbm_server_query(self._h, C.byref(q), C.byref(r), C.byref(f),
                                           C.byref(used), C.byref(budget)))
        return {"queued": q.value, "running": r.value, "finished": f.value,
                "used": used.value, "budget": budget.value}
