"""GPU checks of the on-device reductions, padded plane views, reference
statistics and moments against the oracle / numpy on the same state, and
size-independent properties of the step at the full 256x256x1024 and
256^3 sizes."""
import math

import numpy as np
import pytest

from flowcontrast_amd import KERNEL_PLAIN, KERNEL_TILED, RunParams, Solver, inputs
from flowcontrast_amd._lib import FslbmError

pytestmark = pytest.mark.gpu


def small_pipe(precision="f64"):
    nx, ny, nz, R = 24, 22, 18, 9.0
    mask, _ = inputs.circle_mask(nx, ny, nz, R)
    pin = 3e-3
    st = inputs.poiseuille_state(mask, R, 0.05, pin, 0.05, 3)
    prm = RunParams(tblend=0.9, pin=pin, nu=0.02, d1=2, delta=1, blend=0.5)
    s = Solver(mask.shape, prm, mask=mask, precision=precision)
    s.set_state(*st)
    return s, mask, st, prm


def padded_view(st, mask, d2, pin):
    nz = st[0].shape[0]
    lz = nz + 2 * d2
    zs = (np.arange(lz) - d2) % nz
    jump = np.where(np.arange(lz) < d2, pin, np.where(np.arange(lz) >= lz - d2, -pin, 0.0))
    out = [np.asarray(s, dtype=np.float64)[zs].copy() for s in st]
    out[0] = out[0] + jump[:, None, None]
    wall = mask[zs] != 0
    for f in range(4):
        out[f][wall] = 0.0
    return out


def test_reduce_matches_numpy():
    s, mask, st, prm = small_pipe()
    with s:
        s.step(3)
        got = s.get_state()
        r = s.reduce()
    fl = mask == 0
    assert r[0] == fl.sum()
    assert abs(r[1] - got[0][fl].sum()) < 1e-12
    for f in (1, 2, 3):
        assert abs(r[1 + f] - got[f][fl].sum()) < 1e-13
    assert abs(r[5] - sum((got[f][fl] ** 2).sum() for f in (1, 2, 3))) < 1e-13
    assert abs(r[6] - (got[0][fl] ** 2).sum()) < 1e-12
    assert r[7] == 0


def test_nan_is_detected():
    s, mask, st, prm = small_pipe()
    with s:
        bad = [a.copy() for a in st]
        i = np.argwhere(mask == 0)[5]
        bad[2][tuple(i)] = np.nan
        s.set_state(*bad)
        assert s.reduce()[7] == 1
        s.step(1)
        assert s.reduce()[7] > 1  # spreads to the neighbours


def test_plane_views_and_probe(orc):
    s, mask, st, prm = small_pipe()
    with s:
        s.step(2)
        got = s.get_state()
        d2 = 4
        pv = padded_view(got, mask, d2, prm.pin)
        nz, ny, nx = mask.shape
        for f in range(4):
            yz = s.plane(0, nx // 2, d2, f)
            assert np.array_equal(yz, pv[f][:, :, nx // 2].reshape(-1))
            xy = s.plane(2, d2 + nz // 2, d2, f)
            assert np.array_equal(xy, pv[f][d2 + nz // 2].reshape(-1))
        pr = s.probe(nx // 2, ny // 2, nz // 2)
        assert np.array_equal(pr, [got[f][nz // 2, ny // 2, nx // 2] for f in range(4)])


def test_reference_stats_match_oracle(orc):
    s, mask, st, prm = small_pipe()
    d2 = 3
    nz, ny, nx = mask.shape
    xplane = nx // 2
    o = orc.Stats(ny * (nz + 2 * d2))
    with s:
        s.stats_config(xplane, d2)
        for k, delta in enumerate([2, 2, 1, 2, 2, 0, 2]):
            s.step(5)
            s.stats_accumulate(delta)
            got = s.get_state()
            pv = padded_view(got, mask, d2, prm.pin)
            o.accumulate([pv[f][:, :, xplane].reshape(-1) for f in range(4)], delta)
        acc, cnt, cnt2 = s.stats_read()
    ref, rc, rc2 = o.read()
    assert (cnt, cnt2) == (rc, rc2)
    assert np.abs(acc - ref).max() <= 1e-14 * max(1.0, np.abs(ref).max())


def test_stats_load_roundtrip():
    s, mask, st, prm = small_pipe()
    with s:
        s.stats_config(5, 2)
        s.stats_accumulate(2)
        s.step(2)
        s.stats_accumulate(2)
        acc, cnt, cnt2 = s.stats_read()
        s.stats_config(5, 2)
        s.stats_load(acc, cnt, cnt2)
        acc2, c, c2 = s.stats_read()
    assert np.array_equal(acc, acc2) and (c, c2) == (cnt, cnt2)


def test_moments_match_numpy():
    s, mask, st, prm = small_pipe()
    with s:
        s.moments_reset()
        tot = np.zeros((11,) + mask.shape[1:])
        for _ in range(3):
            s.step(4)
            s.moments_accumulate()
            g = s.get_state()
            p, u, v, w = [np.where(mask == 0, a, 0.0) for a in g]
            for k, q in enumerate([p, u, v, w, p * p, u * u, v * v, w * w, u * v, u * w, v * w]):
                tot[k] += q.sum(axis=0)
        mom, n = s.moments_read()
    assert n == 3
    assert np.abs(mom - tot).max() < 1e-12


@pytest.mark.parametrize("precision", ["f16s", "bf16s"])
def test_shifted_storage_roundtrip(precision):
    s, mask, st, prm = small_pipe(precision)
    with s:
        got = s.get_state()
    # velocities are stored unscaled: each value rounds to nearest with a
    # relative error of half an ulp (2^-11 for FP16, 2^-8 for BF16), or an
    # absolute 2^-25 among FP16 subnormals.  uz is stored as its departure
    # from the column's streamwise reference (the z-mean of uz on a column
    # fluid in every plane, held in float): half an ulp of the departure,
    # plus the float rounding of the decoded sum and of the reference.
    rel = 2 ** -11 if precision == "f16s" else 2 ** -8
    fl = mask == 0
    ref = np.where(fl.all(axis=0), st[3].mean(axis=0), 0.0)[None]
    for f in (1, 2):
        err = np.abs(got[f] - st[f])
        assert (err <= rel * np.abs(st[f]) + 2.0 ** -25).all(), f
    err = np.abs(got[3] - st[3])
    bound = (rel * np.abs(st[3] - ref) + 2.0 ** -24 * (np.abs(st[3]) + 2 * np.abs(ref))
             + 2.0 ** -25)
    assert (err <= bound).all()
    assert np.abs(got[0] - st[0]).max() <= 1e-5
    assert (got[0][mask != 0] == 0).all()


# ----------------------------------------------------------------------
# size-independent properties at the benchmark sizes
# ----------------------------------------------------------------------
def _variants():
    return [KERNEL_PLAIN, KERNEL_TILED]


@pytest.mark.parametrize("variant", _variants())
def test_full_size_shear_wave_is_exact_eigenmode(variant):
    """256^3 periodic box, FP32: u = A sin(kz) x-hat decays by exactly
    lambda(k) per step (the update is linear on this subspace)."""
    n = 256
    k = 2 * math.pi * 4 / n
    tb, nu, b, d1, A = 1.0, 0.01, 1.0, 2, 0.05
    prm = RunParams(tblend=tb, pin=0.0, nu=nu, d1=d1, delta=1, blend=b)
    z = np.arange(n, dtype=np.float64)
    prof = (A * np.sin(k * z))[:, None, None] * np.ones((1, n, n))
    p = np.full((n, n, n), 0.25, dtype=np.float32)
    zero = np.zeros((n, n, n), dtype=np.float32)
    try:
        s = Solver((n, n, n), prm, mask=None, precision="f32")
        s.set_kernel(variant)
    except FslbmError:
        pytest.skip("variant unsupported")
    with s:
        s.set_state(p, prof.astype(np.float32), zero, zero)
        s.step(10)
        got = s.get_state(np.float32)
    c1, c2 = 1 + b / 3, -b / 12
    lam = (1 - tb) + tb * ((2 + math.cos(k)) / 3 + (nu - 1 / 6) / d1 ** 2 *
                           (-2 * c1 * (1 - math.cos(k * d1)) - 2 * c2 * (1 - math.cos(2 * k * d1))))
    exp = lam ** 10 * prof
    assert np.abs(got[1] - exp).max() < 1e-6 * A
    assert np.abs(got[0] - 0.25).max() < 1e-6
    assert np.abs(got[2]).max() < 1e-7 and np.abs(got[3]).max() < 1e-7


# Relative size of one storage ulp at the top of the scaled range (values
# are scaled into [512, 1024)): FP32 has no scaling; FP16 has 11
# significant bits (ulp 2^-10 relative at the top), BF16 8 (2^-7).
STORE_ULP = {"f32": 2.0 ** -23, "f16s": 2.0 ** -10, "bf16s": 2.0 ** -7}


@pytest.mark.parametrize("precision", ["f32", "f16s", "bf16s"])
@pytest.mark.parametrize("variant", _variants())
def test_full_size_pipe_conserves_subspace(variant, precision):
    """256x256x1024 pipe: a z-invariant axial flow with the linear pressure
    ramp stays axial (ux = uy = 0, p unchanged); the rest state is fixed
    (size-independent properties of the update), in FP32 and in the 16-bit
    storage modes.  Tolerances: FP32 roundoff of sums of 27 terms of size
    ~p (0.01) and ~|w| (0.05); in 16-bit storage, the stochastic rounding
    of the stored axial velocity differs from plane to plane by up to two
    storage ulps of |w|, and the cross-stream flow and pressure it drives
    stay within two (velocity) and four (pressure) such ulps."""
    nx = ny = 256
    nz = 1024
    R = 120.0
    mask, nf = inputs.circle_mask(nx, ny, nz, R)
    pin = 1e-3
    G = pin / nz
    fl = mask == 0
    zc = np.arange(nz, dtype=np.float32)[:, None, None]
    p = np.where(fl, np.float32(0.01) - np.float32(G) * zc, 0).astype(np.float32)
    yy, xx = np.mgrid[0:ny, 0:nx]
    r2 = ((xx - 127.5) ** 2 + (yy - 127.5) ** 2) / R ** 2
    w2 = (-0.05 * np.clip(1 - r2, 0, None)).astype(np.float32)
    w = np.where(fl, w2[None], 0).astype(np.float32)
    zero = np.zeros_like(p)
    prm = RunParams(tblend=1.0, pin=pin, nu=0.01, d1=2, delta=1, blend=0.0)
    try:
        s = Solver(mask.shape, prm, mask=mask, precision=precision)
        s.set_kernel(variant)
    except FslbmError:
        pytest.skip("variant unsupported")
    with s:
        s.set_state(p, zero, zero, w)
        s.step(4)
        got = s.get_state(np.float32)
        r = s.reduce()
    assert r[7] == 0
    ulp = STORE_ULP[precision]
    # In 16-bit storage the rounding of the stored axial velocity differs
    # from plane to plane by about an ulp of |w|; that z-variation drives
    # cross-stream velocity and a pressure response of the same size.
    cross = 1e-6 * 0.05 if precision == "f32" else 2 * ulp * 0.05
    ptol = 2e-8 if precision == "f32" else 4 * ulp * 0.05
    assert np.abs(got[1]).max() < cross and np.abs(got[2]).max() < cross
    assert np.abs(got[0] - p).max() < ptol
    # z-invariance of the axial velocity is kept up to roundoff (FP32) or
    # two storage ulps of |w| (16-bit storage)
    tol_z = 1e-6 * 0.05 if precision == "f32" else 2 * ulp * 0.05
    assert np.abs(got[3] - got[3][:1]).max() < tol_z


def test_autotune_keeps_state_and_results():
    """The autotuner times trial launches without advancing the state, and
    the chunk it picks changes only the schedule: results are bitwise the
    same as with the default schedule."""
    nx, ny, nz, R = 70, 60, 96, 26.0
    mask, _ = inputs.circle_mask(nx, ny, nz, R)
    st = inputs.poiseuille_state(mask, R, 0.04, 1e-3, 0.05, 9)
    prm = RunParams(tblend=0.8, pin=1e-3, nu=0.01, d1=2, delta=1, blend=0.5)
    outs = []
    for tune in (False, True):
        with Solver(mask.shape, prm, mask=mask, precision="f32") as s:
            s.set_state(*[a.astype(np.float32) for a in st])
            if tune:
                before = s.get_state(np.float32)
                chunk, ms = s.autotune(2)
                assert 1 <= chunk <= nz and ms > 0
                after = s.get_state(np.float32)
                for a, b in zip(before, after):
                    assert np.array_equal(a, b)
            s.step(9)
            outs.append(s.get_state(np.float32))
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("precision", ["f64", "f32", "f16s", "bf16s"])
def test_snapshot_resumes_bitwise(precision):
    """A snapshot taken mid-run and loaded into a fresh context continues
    bitwise as the original; a state set from doubles does not resume the
    16-bit modes exactly (re-quantised, rounding counter restarted)."""
    nx, ny, nz, R = 40, 36, 24, 15.0
    mask, _ = inputs.circle_mask(nx, ny, nz, R)
    st = inputs.poiseuille_state(mask, R, 0.05, 2e-3, 0.05, 6)
    prm = RunParams(tblend=0.9, pin=2e-3, nu=0.01, d1=2, delta=1, blend=0.0)
    with Solver(mask.shape, prm, mask=mask, precision=precision) as a:
        a.set_state(*st)
        a.step(7)
        a.set_gauge(1e-4)
        snap = a.snapshot()
        a.step(5)
        ref = a.get_state()
    with Solver(mask.shape, prm, mask=mask, precision=precision) as b:
        b.set_state(*st)
        b.load_snapshot(snap)
        b.step(5)
        got = b.get_state()
        for g, e in zip(got, ref):
            assert np.array_equal(g, e)
        with pytest.raises(FslbmError):  # truncated
            b.load_snapshot(snap[:-1])
    with Solver(mask.shape, prm, mask=mask, precision="f32" if precision != "f32" else "f64") as c:
        with pytest.raises(FslbmError):  # another precision mode
            c.load_snapshot(snap)


def test_f16s_tiles_follow_the_creation_blend():
    """FP16 storage in a walled box plans 12-row tiles at blend 0 and 8-row
    tiles otherwise; a context keeps its tiles when its blend changes.  The
    tile shape changes only the schedule: a context created at blend 0 and
    switched to 0.5 gives bitwise the state of one created at 0.5."""
    from flowcontrast_amd import tile_plan
    nx, ny, nz, R = 70, 60, 40, 26.0
    mask, _ = inputs.circle_mask(nx, ny, nz, R)
    st = [a.astype(np.float32) for a in inputs.poiseuille_state(mask, R, 0.04, 1e-3, 0.05, 4)]
    prm0 = RunParams(tblend=0.9, pin=1e-3, nu=0.01, d1=2, delta=1, blend=0.0)
    prm5 = RunParams(tblend=0.9, pin=1e-3, nu=0.01, d1=2, delta=1, blend=0.5)
    assert tile_plan(mask.shape, "f16s", mask, 0.0)["tile_rows"] == 12
    assert tile_plan(mask.shape, "f16s", mask, 0.5)["tile_rows"] == 8
    outs = []
    for first in (prm0, prm5):
        with Solver(mask.shape, first, mask=mask, precision="f16s", kernel=KERNEL_TILED) as s:
            assert s.step_cells == tile_plan(mask.shape, "f16s", mask, first.blend)["cells"]
            s.set_state(*st)
            s.set_params(prm5)
            s.step(6)
            outs.append(s.get_state(np.float32))
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("precision", ["f16s", "bf16s"])
@pytest.mark.parametrize("variant", _variants())
def test_shifted_storage_rounding_is_unbiased(precision, variant):
    """A slowly decaying shear wave in 16-bit storage: its per-step change
    (about 5e-6) is below half a storage ulp of the amplitude 0.05, which
    round-to-nearest would lose every step (the wave would not decay).
    Stochastic rounding keeps the expected decay: after 1000 steps the
    projection on sin(kz) matches lambda(k)^1000 of the closed form."""
    n = 64
    k = 2 * math.pi / n
    tb, nu, b, d1, A = 1.0, 0.01, 0.0, 2, 0.05
    prm = RunParams(tblend=tb, pin=0.0, nu=nu, d1=d1, delta=1, blend=b)
    z = np.arange(n, dtype=np.float64)
    prof = (A * np.sin(k * z))[:, None, None] * np.ones((1, n, n))
    p = np.full((n, n, n), 1.0)
    zero = np.zeros((n, n, n))
    with Solver((n, n, n), prm, mask=None, precision=precision) as s:
        s.set_kernel(variant)
        s.set_state(p, prof, zero, zero)
        s.step(1000)
        got = s.get_state()
    c1, c2 = 1 + b / 3, -b / 12
    lam = (1 - tb) + tb * ((2 + math.cos(k)) / 3 + (nu - 1 / 6) / d1 ** 2 *
                           (-2 * c1 * (1 - math.cos(k * d1)) - 2 * c2 * (1 - math.cos(2 * k * d1))))
    amp = float(np.sum(got[1] * prof) / np.sum(prof * prof))
    assert 1 - lam ** 1000 > 0.08  # the decay to be resolved
    assert abs(amp - lam ** 1000) < 2e-3 * lam ** 1000, (amp, lam ** 1000)


@pytest.mark.parametrize("precision", ["f16s", "bf16s"])
def test_16bit_storage_from_a_quiet_start(precision):
    """From the reference's fresh start (the pressure ramp with the fluid at
    rest, FSLBM.cu:374-383) a strong drive accelerates a pipe to a laminar
    centreline speed of about 0.17, past |u| = 0.125, the largest velocity
    a scale chosen from a quiet upload could hold.  The 16-bit state stays
    finite and its mean axial velocity follows the FP32 run (tolerance:
    ten storage ulps of the relative error)."""
    nx = ny = 20
    nz, R = 32, 8.0
    mask, _ = inputs.circle_mask(nx, ny, nz, R)
    nu, pin = 0.1, 0.1  # laminar ubar = pin R^2 / (24 nu nz) = 0.083
    fl = mask == 0
    z = np.arange(nz)[:, None, None]
    p = np.where(fl, pin * (nz - z) / nz, 0.0)
    zero = np.zeros_like(p)
    prm = RunParams(tblend=1.0, pin=pin, nu=nu, d1=2, delta=1, blend=0.0)
    out = {}
    for prec in ("f32", precision):
        with Solver(mask.shape, prm, mask=mask, precision=prec) as s:
            s.set_state(p, zero, zero, zero)
            s.step(1500)
            r = s.reduce()
            uz = s.get_state()[3]
        assert r[7] == 0, prec
        out[prec] = (r[4] / r[0], np.abs(uz).max())
    assert out["f32"][1] > 0.15 and out[precision][1] > 0.15
    tol = 10 * STORE_ULP[precision]
    assert abs(out[precision][0] / out["f32"][0] - 1) < tol, out
