"""The streamwise reference profile of the 16-bit storage modes.

FP16S and BF16S store uz as its departure from a per-column reference
profile uref(x, y), chosen at upload (the z-mean of uz on a column that is
fluid in every plane, 0 on a column holding a wall cell).  A stored value
keeps 11 (FP16) or 8 (BF16) significant bits of that departure instead of
uz itself, so the velocity rounding of every store, and the pressure noise
it drives through the predictor's e.u terms, shrink with it.

Checked through the C ABI:
  upload       the round-to-nearest upload of a Poiseuille state decodes
               within half an ulp of |uz - uref| (without the reference the
               error would be half an ulp of |uz|, 10x and more larger)
  laminar pipe 1500 steps of a laminar 34x34x96 pipe (Re 3000, 2% noise)
               against FP64: the pressure deviation's relative L2 error.
               An emulation of FP16 storage (FP32 oracle steps with
               stochastically rounded stores) gave 11% without the
               reference and 2.1% with it; the bounds are 5% (FP16) and
               20% (BF16, 3 bits fewer)
  walls        columns holding a wall cell carry no reference: wall cells
               decode to 0 in z-varying masks
"""
import numpy as np
import pytest

from flowcontrast_amd import RunParams, Solver, inputs

pytestmark = pytest.mark.gpu

ULP_HALF = {"f16s": 2.0 ** -11, "bf16s": 2.0 ** -8}


def laminar_pipe():
    nx = ny = 34
    nz, R = 96, 14.0
    re, ubar = 3000.0, 0.05
    nu = ubar * 2 * R / re
    pin = inputs.laminar_pin(nu, nz, ubar, R)
    mask, _ = inputs.circle_mask(nx, ny, nz, R)
    st = inputs.poiseuille_state(mask, R, ubar, pin, 0.02, 3)
    prm = RunParams(tblend=1.0, pin=pin, nu=nu, d1=2, delta=1, blend=0.0)
    return mask, st, prm


@pytest.mark.parametrize("precision", ["f16s", "bf16s"])
def test_upload_rounds_the_departure_from_the_reference(precision):
    mask, st, prm = laminar_pipe()
    fl = mask == 0
    ref = np.where(fl.all(axis=0), st[3].mean(axis=0), 0.0)[None]
    dev = np.abs(np.where(fl, st[3] - ref, 0.0))
    with Solver(mask.shape, prm, mask=mask, precision=precision) as s:
        s.set_state(*st)
        got = s.get_state()
    err = np.abs(got[3] - st[3])
    # half an ulp of the departure (binade of each value), plus the float
    # rounding of the decoded sum uz = stored + uref
    # (and of FP16's subnormal spacing 2^-24 below 2^-14)
    bound = ULP_HALF[precision] * 2.0 ** np.ceil(np.log2(np.maximum(dev, 1e-30))) \
        + 2.0 ** -24 * np.abs(st[3]) + (2.0 ** -25 if precision == "f16s" else 0.0) + 1e-12
    assert np.all(err <= bound), float((err / bound).max())
    # the departures are a small fraction of uz
    assert dev.max() < 0.1 * np.abs(st[3]).max()
    assert not np.any(got[3][~fl])


@pytest.mark.parametrize("precision, bound", [("f16s", 0.05), ("bf16s", 0.20)])
def test_laminar_pipe_pressure_accuracy(precision, bound):
    mask, st, prm = laminar_pipe()
    fl = mask == 0
    with Solver(mask.shape, prm, mask=mask, precision="f64") as s:
        s.set_state(*st)
        s.step(1500)
        ref = s.get_state()
    with Solver(mask.shape, prm, mask=mask, precision=precision) as s:
        s.set_state(*st)
        s.step(1500)
        got = s.get_state()
    rp = ref[0][fl] - ref[0][fl].mean()
    gp = got[0][fl] - got[0][fl].mean()
    ep = float(np.linalg.norm(gp - rp) / np.linalg.norm(rp))
    eu = float(np.sqrt(sum(np.sum((got[f][fl] - ref[f][fl]) ** 2) for f in (1, 2, 3))
                       / sum(np.sum(ref[f][fl] ** 2) for f in (1, 2, 3))))
    print(f"\n{precision}: pressure deviation rel L2 {ep:.3e}, velocity rel L2 {eu:.3e}")
    assert ep <= bound, ep
    assert eu <= bound / 20, eu


def test_walled_columns_carry_no_reference():
    """A pipe with an orifice plate: columns crossing the plate hold wall
    cells in some planes; all wall cells decode to 0 after steps in both
    16-bit modes."""
    nx = ny = 30
    nz = 32
    z, y, x = np.mgrid[0:nz, 0:ny, 0:nx].astype(np.float64)
    r = np.hypot(x - (nx - 1) / 2, y - (ny - 1) / 2)
    wall = (r > 12.0) | ((z >= 10) & (z < 13) & (r > 6.0))
    mask = np.where(wall, 125, 0).astype(np.uint8)
    fl = ~wall
    rng = np.random.default_rng(2)
    st = [np.where(fl, 1e-3 * (nz - z) / nz, 0.0)]
    for base in (0.0, 0.0, -0.04):
        st.append(np.where(fl, base + 0.002 * rng.uniform(-1, 1, mask.shape), 0.0))
    prm = RunParams(tblend=1.0, pin=1e-3, nu=0.02, d1=2, delta=1, blend=0.0)
    for precision in ("f16s", "bf16s"):
        with Solver(mask.shape, prm, mask=mask, precision=precision) as s:
            s.set_state(*st)
            s.step(5)
            got = s.get_state()
            assert s.reduce()[7] == 0
        for f in range(4):
            assert not np.any(got[f][wall]), (precision, f)
