"""fslbm_step_noalloc, the entry point that allocates nothing, against the
context path (fslbm_create / fslbm_set_state / fslbm_step) on the same
inputs: bitwise equal results, for two calls in a row of which the second
has a smaller box (the shared wall-weight table of the first must not leak
into the second), in FP32 and FP64, walled and periodic, and over several
steps ping-ponging between the caller's two buffers.  The caller's device
buffers are allocated here through the HIP runtime the library uses."""
import ctypes as C

import numpy as np
import pytest

from flowcontrast_amd import KERNEL_TILED, RunParams, Solver, inputs
from flowcontrast_amd import _lib

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip():
    L = _lib.lib()  # loads the runtime libfslbm.so links
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipDeviceSynchronize.argtypes = []
    for f in (h.hipMalloc, h.hipMemcpy, h.hipFree, h.hipDeviceSynchronize):
        f.restype = C.c_int
    assert L is not None
    return h


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), nbytes) == 0
        self.nbytes = nbytes

    def put(self, arr):
        a = np.ascontiguousarray(arr)
        assert a.nbytes == self.nbytes
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, dtype, shape):
        a = np.empty(shape, dtype)
        assert self.hip.hipMemcpy(a.ctypes.data, self.ptr, a.nbytes, D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.ptr)


def pipe(nx, ny, nz, R, seed):
    mask, _ = inputs.circle_mask(nx, ny, nz, R)
    st = inputs.poiseuille_state(mask, R, 0.04, 2e-3, 0.05, seed)
    return mask, st


def tgv(n):
    return None, inputs.taylor_green_state(n, 0.05, 1.0)


def context_run(mask, st, prm, precision, steps, gauge):
    shape = np.shape(st[0])
    with Solver(shape, prm, mask=mask, precision=precision, kernel=KERNEL_TILED) as s:
        s.set_state(*st)
        s.set_gauge(gauge)
        s.step(steps)
        return s.get_state()


def noalloc_run(hip, mask, st, prm, precision, steps, gauge):
    dtype = np.float64 if precision == "f64" else np.float32
    nz, ny, nx = np.shape(st[0])
    state = np.stack([np.asarray(a, dtype) for a in st])
    a, b = DevBuf(hip, state.nbytes), DevBuf(hip, state.nbytes)
    m = None
    try:
        a.put(state)
        if mask is not None:
            m = DevBuf(hip, nx * ny)
            m.put(np.ascontiguousarray(mask[0], dtype=np.uint8))
        L = _lib.lib()
        cp = prm.to_c()
        for k in range(steps):
            rc = L.fslbm_step_noalloc(C.byref(cp), m.ptr if m else None, nx, ny, nz,
                                      _lib.PRECISIONS[precision], a.ptr, b.ptr,
                                      gauge if k == 0 else 0.0, None)
            assert rc == 0, L.fslbm_last_error(None)
            a, b = b, a
        assert hip.hipDeviceSynchronize() == 0
        out = a.get(dtype, state.shape)
    finally:
        a.free()
        b.free()
        if m:
            m.free()
    return [out[f].astype(np.float64) for f in range(4)]


CASES = [
    # (case, params, steps): the second of each pair is smaller
    (lambda: pipe(70, 60, 40, 26.0, 1), RunParams(0.9, 2e-3, 0.02, 2, 1, 0.0), 1),
    (lambda: pipe(40, 36, 24, 15.0, 2), RunParams(1.0, 2e-3, 0.015, 1, 1, 0.5), 1),
    (lambda: pipe(131, 77, 19, 37.0, 3), RunParams(0.9, 2e-3, 0.02, 2, 1, 1.0), 5),
    (lambda: tgv(32), RunParams(1.0, 0.0, 0.01, 2, 1, 0.0), 4),
]


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_noalloc_equals_context_twice(hip, precision):
    for build, prm, steps in CASES:
        mask, st = build()
        if precision == "f32":
            st = [a.astype(np.float32) for a in st]
        ref = context_run(mask, st, prm, precision, steps, 3e-4)
        got = noalloc_run(hip, mask, st, prm, precision, steps, 3e-4)
        for f in range(4):
            assert np.array_equal(got[f], ref[f]), (np.shape(st[0]), precision, f)


def test_noalloc_rejects_what_it_cannot_hold(hip):
    L = _lib.lib()
    cp = RunParams(1.0, 2e-3, 0.02, 2, 1, 0.0).to_c()
    fake = C.c_void_p(0x1000)  # never dereferenced: validation fails first
    side = 1024  # 2^20 columns > FSLBM_NOALLOC_MAX_COLUMNS
    rc = L.fslbm_step_noalloc(C.byref(cp), fake, side, side, 16, _lib.F32, fake,
                              C.c_void_p(0x2000), 0.0, None)
    assert rc == -1 and b"columns" in L.fslbm_last_error(None)
    rc = L.fslbm_step_noalloc(C.byref(cp), fake, 64, 64, 16, _lib.F16S, fake,
                              C.c_void_p(0x2000), 0.0, None)
    assert rc == -1 and b"FSLBM_F32" in L.fslbm_last_error(None)
    rc = L.fslbm_step_noalloc(C.byref(cp), fake, 64, 64, 16, _lib.F32, fake, fake, 0.0, None)
    assert rc == -1 and b"different" in L.fslbm_last_error(None)
    # a periodic box needs no table: any column count is fine (not run here)
    assert _lib.NOALLOC_MAX_COLUMNS == 1 << 19
