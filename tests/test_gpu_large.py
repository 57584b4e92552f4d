"""Domains beyond 2^31 field values: the paper's 580x580x1693 grid
(MethodDoc_withFSLBMpaper.pdf p.9, 569.5 million cells; its FP32 state of
four fields is 9.1 GB per copy, each field 2.28 GB, past 2^31 bytes).
Field offsets are 64-bit in the plain and FP64 kernels; the pipelined
kernel addresses each field (FP32) or pair plane (16-bit storage) of up to
4 GiB through a buffer resource.

Size-independent properties, checked after a few steps through the C ABI:
  rest state       p = const, u = 0 on the fluid: a fixed point
  Poiseuille       z-invariant axial flow with the linear pressure ramp:
                   stays axial, pressure unchanged, axial flow z-invariant
Tolerances as for the 256x256x1024 pipe (test_gpu_aux.py): FP32 roundoff of
sums of 27 terms of size ~p and ~|w|; in 16-bit storage two storage ulps of
|w| for the plane-to-plane rounding and what it drives.
"""
import numpy as np
import pytest

from flowcontrast_amd import KERNEL_PLAIN, KERNEL_TILED, RunParams, Solver, inputs, tile_plan

pytestmark = pytest.mark.gpu

NX = NY = 580
NZ = 1693
R = 280.0
STORE_ULP = {"f32": 2.0 ** -23, "bf16s": 2.0 ** -7}


@pytest.fixture(scope="module")
def big():
    mask, nf = inputs.circle_mask(NX, NY, NZ, R)
    assert mask.size == 580 * 580 * 1693 and 4 * mask.size > 2 ** 31
    fl = mask[0] == 0
    yy, xx = np.mgrid[0:NY, 0:NX]
    r2 = ((xx - (NX - 1) / 2) ** 2 + (yy - (NY - 1) / 2) ** 2) / R ** 2
    w2 = np.where(fl, -0.05 * np.clip(1 - r2, 0, None), 0).astype(np.float32)
    return mask, fl, w2


@pytest.mark.parametrize("precision, variant", [("f32", KERNEL_TILED), ("f32", KERNEL_PLAIN),
                                                ("bf16s", KERNEL_TILED), ("f64", KERNEL_TILED)])
def test_paper_grid_properties(big, precision, variant):
    """FP64 at this size is past the pipelined kernel's 32-bit buffer
    offsets (8 bytes x 569 M cells per field) and runs the two-barrier
    kernel on 8-row tiles; its results are checked as float32 arrays, at
    the FP32 tolerances."""
    mask, fl, w2 = big
    if precision == "f64":
        assert tile_plan(mask.shape, "f64", mask[:1])["tile_rows"] == 8
    pin = 1e-3
    prm = RunParams(tblend=1.0, pin=pin, nu=0.01, d1=2, delta=1, blend=0.0)
    rest = RunParams(tblend=1.0, pin=0.0, nu=0.01, d1=2, delta=1, blend=0.0)
    with Solver(mask.shape, rest, mask=mask, precision=precision) as s:
        s.set_kernel(variant)
        assert s.step_cells <= mask.size
        # rest state (no pressure jump): a fixed point
        p0 = np.where(mask == 0, np.float32(0.37), np.float32(0))
        zero = np.zeros(mask.shape, np.float32)
        s.set_state(p0, zero, zero, zero)
        s.step(2)
        got = s.get_state(np.float32)
        assert s.reduce()[7] == 0
        tol_p = 2e-7 if precision in ("f32", "f64") else 4 * STORE_ULP[precision] * 0.37
        assert np.abs(got[0] - p0).max() < tol_p
        for f in (1, 2, 3):
            assert np.abs(got[f]).max() < 1e-7, f
        del got
        # Poiseuille subspace, driven by the jump pin
        s.set_params(prm)
        G = np.float32(pin / NZ)
        zc = np.arange(NZ, dtype=np.float32)[:, None, None]
        p = np.where(mask == 0, np.float32(0.01) - G * zc, np.float32(0))
        w = np.broadcast_to(w2[None], mask.shape)
        s.set_state(p, zero, zero, w)
        s.step(3)
        got = s.get_state(np.float32)
        r = s.reduce()
    assert r[7] == 0
    exact = precision in ("f32", "f64")  # no storage rounding
    ulp = STORE_ULP.get(precision, 0.0)
    cross = 1e-6 * 0.05 if exact else 2 * ulp * 0.05
    # FP32: a sum of 27 terms of size ~|p| = 0.01 rounds by at most
    # 27 * 2^-24 * 0.01 = 1.6e-8 per step; three steps
    ptol = 3 * 27 * 2.0 ** -24 * 0.01 if exact else 4 * ulp * 0.05
    assert np.abs(got[1]).max() < cross and np.abs(got[2]).max() < cross
    assert np.abs(got[0] - p).max() < ptol
    tol_z = 1e-6 * 0.05 if exact else 2 * ulp * 0.05
    assert np.abs(got[3] - got[3][:1]).max() < tol_z
    wall = mask != 0
    for f in range(4):
        assert not np.any(got[f][wall]), f
