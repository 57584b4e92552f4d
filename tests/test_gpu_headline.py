"""GPU parity at the benchmark sizes, against the CPU oracle directly.

The 256x256x1024 pipe (radius 120, bulk Re 5300, Blasius pin, Poiseuille
+ 5% seeded noise: the benchmark's workload), the same lattice with a rough
wall (a mask that varies along z: the tiled kernel's variant for such masks,
the plain kernel in FP64) and the 256^3 Taylor-Green box (U0 0.05, Re 1600)
are stepped a few steps on the GPU in every precision mode and by the oracle
in FP64 (and FP32), from the same initial state.
The pipe runs apply the reference's gauge (monitor, then the gauge
subtracted at the next step, FSLBM.cu:628-647) after step 2.

Tolerances (after 4 steps):
  FP64                 max |gpu - oracle| / max |oracle| <= 1e-12 for p and
                       for the velocity vector (components normalised by the
                       largest velocity component: the Taylor-Green uz starts
                       at zero and is still tiny after 4 steps)
  FP32                 relative L2 <= 1e-5 on p and on the velocity vector,
                       against the FP32 and the FP64 oracle
  FP16S / BF16S        relative L2 against the FP64 oracle of the velocity
                       vector and of the pressure's departure from its mean
                       (TOL16 below, about 3x the measured values).  The
                       state is stored in 16 bits (11 / 8 significant bits)
                       and rounded again at every step, so BF16S errors are
                       8x FP16S ones.  The pipe stores uz as its departure
                       from the column's streamwise reference, hence its
                       small velocity error; the Taylor-Green pressure
                       departure is second order in U0 while the velocity
                       rounding feeds it at first order.
"""
import numpy as np
import pytest

from flowcontrast_amd import KERNEL_AUTO, RunParams, Solver, inputs

pytestmark = pytest.mark.gpu

STEPS = 4
GAUGE_AFTER = 2
# (velocity, pressure departure) bounds; measured after 4 steps:
#   pipe FP16S 9.7e-6, 5.3e-4; BF16S 7.8e-5, 4.2e-3
#   rough pipe FP16S 1.7e-5, 6.9e-4; BF16S 1.3e-4, 5.5e-3
#   Taylor-Green FP16S 3.6e-4, 3.4e-3; BF16S 2.8e-3, 2.7e-2
TOL16 = {("pipe", "f16s"): (3e-5, 2e-3), ("pipe", "bf16s"): (2.5e-4, 1.5e-2),
         ("rough", "f16s"): (5e-5, 2e-3), ("rough", "bf16s"): (4e-4, 1.6e-2),
         ("taylor_green", "f16s"): (1e-3, 1e-2), ("taylor_green", "bf16s"): (8e-3, 8e-2)}


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def vel_l2(got, ref):
    num = sum(float(np.sum((np.asarray(got[f], dtype=np.float64) - ref[f]) ** 2))
              for f in (1, 2, 3))
    den = sum(float(np.sum(np.asarray(ref[f], dtype=np.float64) ** 2)) for f in (1, 2, 3))
    return float(np.sqrt(num / den))


def dev_l2(got, ref, mask):
    """Relative L2 of the pressure's departure from its fluid mean."""
    fl = mask == 0 if mask is not None else np.ones(np.shape(ref), bool)
    r = np.asarray(ref)[fl]
    g = np.asarray(got, dtype=np.float64)[fl]
    return rel_l2(g - r.mean(), r - r.mean())


def max_rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


def vel_max(got, ref):
    umax = max(float(np.abs(ref[f]).max()) for f in (1, 2, 3))
    return max(float(np.abs(np.asarray(got[f]) - ref[f]).max()) for f in (1, 2, 3)) / umax


def pipe_case():
    nx = ny = 256
    nz = 1024
    R, re = 120.0, 5300.0
    nu = 0.05 * 2 * R / re
    mask, _ = inputs.circle_mask(nx, ny, nz, R)
    pin = inputs.blasius_pin(nu, nz, re, R)
    ubar = re * nu / (2 * R)
    st = inputs.poiseuille_state(mask, R, ubar, pin, 0.05, 3)
    prm = RunParams(tblend=1.0, pin=pin, nu=nu, d1=2, delta=1, blend=0.0)
    return mask, st, prm


def rough_case():
    """The pipe's lattice with a rough wall (a mask that varies along z):
    mean radius 120, corrugation of 4 cells and 600 seeded bumps."""
    nx = ny = 256
    nz, R, amp = 1024, 120.0, 4.0
    re = 5300.0
    nu = 0.05 * 2 * R / re
    pin = inputs.blasius_pin(nu, nz, re, R)
    mask, _ = inputs.rough_pipe_mask(nx, ny, nz, R, amp, 600, 5)
    st = inputs.poiseuille_state(mask, R, re * nu / (2 * R), pin, 0.05, 3)
    prm = RunParams(tblend=1.0, pin=pin, nu=nu, d1=2, delta=1, blend=0.0)
    return mask, st, prm


def tgv_case():
    n = 256
    nu = inputs.tgv_viscosity(n, 0.05, 1600.0)
    st = inputs.taylor_green_state(n, 0.05, 1.0)
    prm = RunParams(tblend=1.0, pin=0.0, nu=nu, d1=2, delta=1, blend=0.0)
    return None, st, prm


def oracle_run(orc, mask, st, prm, dtype):
    """STEPS oracle steps; the pipe's gauge p1 - pin (mean pressure of the
    first plane's fluid cells) is taken after GAUGE_AFTER steps and
    subtracted at the next one."""
    op = orc.params(prm.tblend, prm.pin, prm.nu, prm.blend, prm.d1)
    ref = orc.step_folded(st, op, mask=mask, nsteps=GAUGE_AFTER, dtype=dtype)
    gauge = 0.0
    if mask is not None:
        fl0 = mask[0] == 0
        gauge = float(np.asarray(ref[0][0], dtype=np.float64)[fl0].sum() / fl0.sum()) - prm.pin
    return orc.step_folded(ref, op, mask=mask, gauge=gauge, nsteps=STEPS - GAUGE_AFTER,
                           dtype=dtype)


def gpu_run(mask, st, prm, precision):
    shape = np.shape(st[0])
    with Solver(shape, prm, mask=mask, precision=precision, kernel=KERNEL_AUTO) as s:
        s.set_state(*(st if precision == "f64" else [a.astype(np.float32) for a in st]))
        s.step(GAUGE_AFTER)
        if mask is not None:
            s.monitor(set_gauge=True)
        s.step(STEPS - GAUGE_AFTER)
        return s.get_state()


@pytest.fixture(scope="module", params=["pipe", "rough", "taylor_green"])
def case(request, orc):
    mask, st, prm = {"pipe": pipe_case, "rough": rough_case,
                     "taylor_green": tgv_case}[request.param]()
    ref64 = oracle_run(orc, mask, st, prm, np.float64)
    ref32 = oracle_run(orc, mask, st, prm, np.float32)
    return request.param, mask, st, prm, ref64, ref32


@pytest.mark.parametrize("precision", ["f64", "f32", "f16s", "bf16s"])
def test_benchmark_sizes_match_oracle(case, precision):
    name, mask, st, prm, ref64, ref32 = case
    got = gpu_run(mask, st, prm, precision)
    errs = {"p_dev_l2": dev_l2(got[0], ref64[0], mask), "u_l2": vel_l2(got, ref64),
            "p_l2": rel_l2(got[0], ref64[0])}
    if precision == "f64":
        errs.update({f"max_rel_{f}": max_rel(got[f], ref64[f]) for f in range(4)})
        errs["u_max"] = vel_max(got, ref64)
    print(f"\n{name} {precision} after {STEPS} steps vs FP64 oracle: "
          + " ".join(f"{k} {v:.3e}" for k, v in errs.items()))
    for a in got:
        assert np.isfinite(a).all()
    if precision == "f64":
        assert errs["max_rel_0"] <= 1e-12 and errs["u_max"] <= 1e-12, (name, errs)
    elif precision == "f32":
        assert errs["p_l2"] <= 1e-5 and errs["u_l2"] <= 1e-5, (name, errs)
        e32 = (rel_l2(got[0], ref32[0]), vel_l2(got, ref32))
        print(f"{name} f32 vs FP32 oracle: p_l2 {e32[0]:.3e} u_l2 {e32[1]:.3e}")
        assert max(e32) <= 1e-5, (name, e32)
    else:
        tol_u, tol_p = TOL16[(name, precision)]
        assert errs["u_l2"] <= tol_u and errs["p_dev_l2"] <= tol_p, (name, precision, errs)
