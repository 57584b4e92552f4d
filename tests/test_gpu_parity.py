"""GPU parity of the HIP step against the CPU oracle, through the C ABI.

Tolerances:
  FP64 GPU vs FP64 oracle   max |gpu - oracle| / max |oracle| <= 1e-12 after
                            100 and 1000 steps, for p and for the velocity
                            vector (every component normalised by the
                            largest velocity component: the noise components
                            ux, uy are ~1% of uz and decay)
  FP32 GPU vs FP64 oracle   relative L2 <= 1e-5 on p and on the velocity
  FP32 GPU vs FP32 oracle   vector (all three components together)
Every kernel variant that supports the case is checked.
"""
import math

import numpy as np
import pytest

from flowcontrast_amd import KERNEL_PLAIN, KERNEL_TILED, RunParams, Solver, inputs, tile_plan
from flowcontrast_amd._lib import FslbmError

pytestmark = pytest.mark.gpu

VARIANTS = [KERNEL_PLAIN, KERNEL_TILED]


def max_rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def errors_max(got, ref):
    """[p error, velocity error], max norms; velocity normalised jointly."""
    ep = max_rel(got[0], ref[0])
    umax = max(np.abs(ref[f]).max() for f in (1, 2, 3))
    eu = max(np.abs(np.asarray(got[f]) - ref[f]).max() for f in (1, 2, 3)) / umax
    return ep, eu


def errors_l2(got, ref):
    """[p error, velocity error], relative L2 (velocity as one vector)."""
    ep = rel_l2(got[0], ref[0])
    num = sum(np.sum((np.asarray(got[f], dtype=np.float64) - ref[f]) ** 2) for f in (1, 2, 3))
    den = sum(np.sum(np.asarray(ref[f], dtype=np.float64) ** 2) for f in (1, 2, 3))
    return ep, float(np.sqrt(num / den))


def c1_case(blend=0.0):
    """32x32x64 pipe, R=14, FP64, Re_b ~ 200 (nu = 0.01, ubar = 0.0714)."""
    nx = ny = 32
    nz = 64
    R, nu = 14.0, 0.01
    ubar = 200 * nu / (2 * R)
    pin = inputs.laminar_pin(nu, nz, ubar, R)
    mask, _ = inputs.circle_mask(nx, ny, nz, R)
    st = inputs.poiseuille_state(mask, R, ubar, pin, 0.01, 1)
    prm = RunParams(tblend=1.0, pin=pin, nu=nu, d1=2, delta=1, blend=blend)
    return mask, st, prm


def orc_params(orc, prm):
    return orc.params(prm.tblend, prm.pin, prm.nu, prm.blend, prm.d1)


def make_solver(shape, prm, mask, precision, variant):
    s = Solver(shape, prm, mask=mask, precision=precision)
    try:
        s.set_kernel(variant)
    except FslbmError:
        s.close()
        pytest.skip("kernel variant does not support this case")
    return s


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("blend", [0.0, 1.0])
def test_c1_pipe_f64_100_steps(orc, variant, blend):
    mask, st, prm = c1_case(blend)
    ref = orc.step_folded(st, orc_params(orc, prm), mask=mask, nsteps=100)
    with make_solver(mask.shape, prm, mask, "f64", variant) as s:
        s.set_state(*st)
        s.step(100)
        got = s.get_state()
    for f in range(4):
        assert max_rel(got[f], ref[f]) <= 1e-12, f
    assert max(errors_max(got, ref)) <= 1e-12


@pytest.mark.parametrize("variant", VARIANTS)
def test_c1_pipe_f64_1000_steps_with_gauge(orc, variant):
    """The reference cadence: monitoring every 50 steps sets the gauge p2
    subtracted at the next step (FSLBM.cu:628-647)."""
    mask, st, prm = c1_case(0.0)
    op = orc_params(orc, prm)
    ref = [np.array(a) for a in st]
    with make_solver(mask.shape, prm, mask, "f64", variant) as s:
        s.set_state(*st)
        gauge = 0.0
        at100 = None
        for block in range(20):
            ref = orc.step_folded(ref, op, mask=mask, gauge=gauge, nsteps=50)
            s.step(50)
            fluid0 = mask[0] == 0
            p1 = ref[0][0][fluid0].sum() / fluid0.sum()
            gauge = p1 - prm.pin
            mon = s.monitor(set_gauge=True)
            assert abs(mon[2] - p1) <= 1e-12 * max(1.0, abs(p1))
            assert abs(mon[3] - gauge) <= 1e-12 * max(1.0, abs(p1))
            assert abs(mon[0] - ref[3][0][fluid0].sum() / fluid0.sum()) < 1e-14
            if block == 1:
                at100 = s.get_state()
                assert max(errors_max(at100, ref)) <= 1e-12
                # per field, each normalised by its own largest value
                for f in range(4):
                    assert max_rel(at100[f], ref[f]) <= 1e-12, (100, f)
        got = s.get_state()
    assert max(errors_max(got, ref)) <= 1e-12
    # Per field after 1000 steps.  The cross-stream components ux, uy start
    # as 1% noise and decay by two orders of magnitude, while the rounding
    # of each step is set by the terms of the whole update (p, uz); per
    # field, normalised by their own (decayed) maxima, they are held to
    # 1e-11.  p and uz stay at 1e-12.
    per_field = [max_rel(got[f], ref[f]) for f in range(4)]
    print(f"\nFP64 1000 steps, kernel {variant}: per-field max rel "
          + " ".join(f"{e:.3e}" for e in per_field))
    assert per_field[0] <= 1e-12 and per_field[3] <= 1e-12, per_field
    assert max(per_field) <= 1e-11, per_field


@pytest.mark.parametrize("variant", VARIANTS)
def test_c1_pipe_f32_vs_f64_oracle(orc, variant):
    mask, st, prm = c1_case(0.0)
    ref = orc.step_folded(st, orc_params(orc, prm), mask=mask, nsteps=100)
    ref32 = orc.step_folded(st, orc_params(orc, prm), mask=mask, nsteps=100,
                            dtype=np.float32)
    with make_solver(mask.shape, prm, mask, "f32", variant) as s:
        s.set_state(*[a.astype(np.float32) for a in st])
        s.step(100)
        got = s.get_state()
    assert max(errors_l2(got, ref)) <= 1e-5
    assert max(errors_l2(got, ref32)) <= 1e-5


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("precision,tol", [("f64", 1e-12), ("f32", 1e-5)])
def test_taylor_green_periodic(orc, variant, precision, tol):
    n = 32
    nu = inputs.tgv_viscosity(n, 0.05, 1600.0)
    st = inputs.taylor_green_state(n, 0.05, 1.0)
    prm = RunParams(tblend=1.0, pin=0.0, nu=nu, d1=2, delta=1, blend=0.0)
    ref = orc.step_folded(st, orc_params(orc, prm), nsteps=100)
    with make_solver((n, n, n), prm, None, precision, variant) as s:
        s.set_state(*st)
        s.step(100)
        got = s.get_state()
    err = errors_max(got, ref) if precision == "f64" else errors_l2(got, ref)
    assert max(err) <= tol, err


ODD_CASES = [(37, 29, 23, 11.0, 1, 0.5),
             (70, 66, 17, 30.0, 3, 1.0),
             (9, 9, 8, 3.0, 2, 0.0),
             (131, 77, 19, 37.0, 2, 1.0)]


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("case", range(len(ODD_CASES)))
def test_odd_shapes_and_d1(orc, variant, case):
    """Ragged sizes (not multiples of the tile), d1 = 1, 2, 3, blend.  The
    tiled kernel takes d1 <= 2 and nz >= 16; other cases skip for it."""
    for (nx, ny, nz, R, d1, blend) in [ODD_CASES[case]]:
        mask, _ = inputs.circle_mask(nx, ny, nz, R)
        pin = 2e-3
        st = inputs.poiseuille_state(mask, R, 0.03, pin, 0.05, 5)
        prm = RunParams(tblend=0.9, pin=pin, nu=0.02, d1=d1, delta=1, blend=blend)
        ref = orc.step_folded(st, orc_params(orc, prm), mask=mask, gauge=1e-3, nsteps=7)
        with make_solver(mask.shape, prm, mask, "f64", variant) as s:
            s.set_state(*st)
            s.set_gauge(1e-3)
            s.step(7)
            got = s.get_state()
        assert max(errors_max(got, ref)) <= 1e-12, (nx, ny, nz)


PIPE_CASES = [(37, 29, 23, 11.0, 1, 0.5),    # d1 = 1, order-4 arms
              (45, 40, 20, 18.0, 1, 0.0),    # d1 = 1, blend 0
              (70, 66, 17, 30.0, 2, 0.0),    # ny = 66: not a multiple of 8 or 12
              (131, 77, 19, 37.0, 2, 1.0),   # two tile columns, ragged rows
              (96, 90, 33, 44.0, 2, 0.25)]   # 90 rows, blend between orders

# 16-bit storage tolerances (relative L2 against the FP64 oracle after
# 7 steps).  A stored value carries a relative error of at most one unit in
# the last place of the significand (stochastic rounding): 2^-10 for FP16,
# 2^-7 for BF16.  Independent errors over 7 stores grow to about
# ulp * sqrt(7/3); the bounds are 3x that.  The pressure is stored as the
# deviation from a mean line pbase + pslope*z, so its error is set by the
# deviation's magnitude; its bound is stated against the velocity scale
# below.
TOL16 = {"f16s": 3 * 2.0 ** -10 * (7 / 3) ** 0.5, "bf16s": 3 * 2.0 ** -7 * (7 / 3) ** 0.5}


@pytest.mark.parametrize("precision", ["f32", "f16s", "bf16s"])
@pytest.mark.parametrize("case", range(len(PIPE_CASES)))
def test_pipelined_kernel_odd_shapes(orc, precision, case):
    """k_step_pipe (the kernel of every FP32 and 16-bit storage run) on
    ragged walled boxes, d1 = 1 and 2, blend 0 and != 0, with a gauge:
    FP32 within relative L2 1e-5 of the FP32 oracle; FP16/BF16 storage
    within the storage tolerance of the FP64 oracle (velocity vector), and
    the pressure error within the same fraction of the velocity scale."""
    nx, ny, nz, R, d1, blend = PIPE_CASES[case]
    mask, _ = inputs.circle_mask(nx, ny, nz, R)
    pin = 2e-3
    st = inputs.poiseuille_state(mask, R, 0.03, pin, 0.05, 5)
    prm = RunParams(tblend=0.9, pin=pin, nu=0.02, d1=d1, delta=1, blend=blend)
    op = orc_params(orc, prm)
    with make_solver(mask.shape, prm, mask, precision, KERNEL_TILED) as s:
        s.set_state(*([a.astype(np.float32) for a in st] if precision == "f32" else st))
        s.set_gauge(1e-3)
        s.step(7)
        got = s.get_state()
    wall = mask != 0
    for f in range(4):
        assert not np.any(np.asarray(got[f])[wall]), f
    if precision == "f32":
        ref32 = orc.step_folded([a.astype(np.float32) for a in st], op, mask=mask,
                                gauge=1e-3, nsteps=7, dtype=np.float32)
        assert max(errors_l2(got, ref32)) <= 1e-5, (nx, ny, nz)
        return
    ref = orc.step_folded(st, op, mask=mask, gauge=1e-3, nsteps=7)
    _, eu = errors_l2(got, ref)
    assert eu <= TOL16[precision], (precision, eu)
    # pressure: absolute rms error against the velocity rms
    urms = np.sqrt(np.mean(sum(np.asarray(ref[f]) ** 2 for f in (1, 2, 3))))
    prms = np.sqrt(np.mean((np.asarray(got[0]) - ref[0]) ** 2))
    assert prms <= TOL16[precision] * urms, (precision, prms / urms)


def off_centre_pipe(nx=200, ny=120, nz=24, cx=40.5, cy=30.5, R=15.0, seed=8):
    """A small pipe in a corner of a wide box: most tiles hold only wall."""
    y, x = np.mgrid[0:ny, 0:nx]
    disc = (x - cx) ** 2 + (y - cy) ** 2 <= R * R
    mask = np.where(np.broadcast_to(disc, (nz, ny, nx)), 0, 125).astype(np.uint8)
    rng = np.random.default_rng(seed)
    fl = mask == 0
    r2 = ((x - cx) ** 2 + (y - cy) ** 2) / (R * R)
    pin, ubar = 2e-3, 0.03
    z = np.arange(nz)[:, None, None]
    p = np.where(fl, pin * (nz - z) / nz, 0.0)
    uz = np.where(fl, -2 * ubar * (1 - r2)[None], 0.0)
    st = [p] + [np.where(fl, u + 0.05 * ubar * rng.uniform(-1, 1, mask.shape), 0.0)
                for u in (np.zeros(mask.shape), np.zeros(mask.shape), uz)]
    return mask, st, pin


@pytest.mark.parametrize("precision, tol", [("f64", 1e-12), ("f32", 1e-5)])
def test_wall_only_tiles_are_skipped_exactly(orc, precision, tol):
    """The tiled kernel runs no work item on tiles of wall cells.  After an
    odd number of steps the state sits in the other buffer: its wall cells
    must still be exactly zero and the fluid must match the oracle."""
    mask, st, pin = off_centre_pipe()
    prm = RunParams(tblend=0.9, pin=pin, nu=0.02, d1=2, delta=1, blend=1.0)
    ref = orc.step_folded(st, orc_params(orc, prm), mask=mask, gauge=5e-4, nsteps=7)
    inp = st if precision == "f64" else [a.astype(np.float32) for a in st]
    with make_solver(mask.shape, prm, mask, precision, KERNEL_TILED) as s:
        s.set_state(*inp)
        s.set_gauge(5e-4)
        s.step(7)
        got = s.get_state()
        plan = tile_plan(mask.shape, precision, mask, prm.blend)
        assert s.step_cells == plan["cells"] < mask.size
        assert s.tile_chunk == plan["chunk"]
    wall = mask != 0
    for f in range(4):
        assert not np.any(np.asarray(got[f])[wall]), f
    err = errors_max(got, ref) if precision == "f64" else errors_l2(got, ref)
    assert max(err) <= tol


def test_plain_and_tiled_agree_f32():
    mask, st, prm = c1_case(1.0)
    outs = []
    for v in VARIANTS:
        try:
            with make_solver(mask.shape, prm, mask, "f32", v) as s:
                s.set_state(*[a.astype(np.float32) for a in st])
                s.step(50)
                outs.append(s.get_state())
        except pytest.skip.Exception:
            pass
    if len(outs) < 2:
        pytest.skip("only one kernel variant available")
    # two FP32 evaluation orders: roundoff only, within the FP32 tolerance
    assert max(errors_l2(outs[0], outs[1])) <= 1e-5


@pytest.mark.parametrize("precision, tol", [("f32", 5e-5), ("f16s", 2e-3), ("bf16s", 1e-2)])
def test_taylor_green_kinetic_energy_decay(precision, tol):
    """64^3 Taylor-Green vortex (U0 0.05, Re 400), 1000 steps: the kinetic
    energy history of each storage mode follows the FP64 run.  FP32 runs
    the reference's float literals (1/6 as 0.16666666f, weights to 9-11
    digits), whose effective viscosity and weight sum differ from the exact
    rationals of FP64 by ~1e-8, on top of 1000 steps of float rounding."""
    n, steps, every = 64, 1000, 100
    nu = inputs.tgv_viscosity(n, 0.05, 400.0)
    prm = RunParams(tblend=1.0, pin=0.0, nu=nu, d1=2, delta=1, blend=0.0)
    st = inputs.taylor_green_state(n, 0.05, 1.0)

    def history(prec):
        with Solver((n, n, n), prm, mask=None, precision=prec) as s:
            s.set_state(*st)
            ke = []
            for t in range(0, steps + 1, every):
                if t:
                    s.step(every)
                ke.append(s.reduce()[5])
            return np.array(ke)

    ref, got = history("f64"), history(precision)
    assert ref[-1] < 0.9 * ref[0]  # the decay to be followed
    assert np.max(np.abs(got - ref) / ref) < tol
