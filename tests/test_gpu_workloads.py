"""The two multi-solve workloads, run in full on the GPU and checked
against solo runs (bitwise) and against the CPU oracle:

* 64 uneven pipes in one batch (radius U[12,48], length U[4,10] diameters,
  Re U[200,3000], Poiseuille + 2% noise, seed 4), FP32, hipGraph replay,
  50 steps: every pipe equals its solo run bitwise, and the three smallest
  match the oracle (velocity within relative L2 1e-5 of the FP32 oracle;
  pressure, see check_f32).
* the seeded serving stream (32 requests, 80% pipes / 20% Taylor-Green
  boxes, FP32 or FP16 storage, seed 5) with step budgets cut 100-fold:
  every request equals its solo run bitwise, the smallest FP32 request
  matches the oracle as the pipes above do, and the smallest FP16-storage
  request within the storage tolerance of the FP64 oracle.

FP16 storage tolerance: a stored value carries a relative rounding error of
at most one unit in the 11-bit significand (2^-10, stochastic rounding), so
after n steps the velocity's relative L2 error is about 2^-10 * sqrt(n / 3)
for independent roundings; 4e-3 bounds it for n <= 200.
"""
import numpy as np
import pytest

from flowcontrast_amd import KERNEL_AUTO, Batch, Server, Solver
from flowcontrast_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel_l2_pu(got, ref):
    def l2(fs):
        num = sum(np.sum((np.asarray(got[f], np.float64) - ref[f]) ** 2) for f in fs)
        den = sum(np.sum(np.asarray(ref[f], np.float64) ** 2) for f in fs)
        return float(np.sqrt(num / max(den, 1e-300)))
    return l2((0,)), l2((1, 2, 3))


def solo(case, prm, mask, st, steps, precision):
    with Solver(case.shape, prm, mask=mask, precision=precision, kernel=KERNEL_AUTO) as s:
        s.set_state(*st)
        s.step(steps)
        return s.get_state()


def oracle_run(orc, prm, mask, st, steps, dtype):
    op = orc.params(prm.tblend, prm.pin, prm.nu, prm.blend, prm.d1)
    return orc.step_folded(st, op, mask=mask, nsteps=steps, dtype=dtype)


def check_f32(orc, got, prm, mask, st32, steps, what):
    """FP32 against the oracle: the velocity within relative L2 1e-5 of the
    FP32 oracle.  The pressure of these pipes carries no background density
    (its rms is a few 1e-4, the pin-sized ramp) while the sums that make it
    hold |u|^2 terms of ~2e-3, so FP32 rounding alone moves it by ~1e-5
    relative: it is held to the FP64 oracle within twice the FP32 oracle's
    own distance from the FP64 oracle (and 1e-5 at least)."""
    orc32 = oracle_run(orc, prm, mask, st32, steps, np.float32)
    orc64 = oracle_run(orc, prm, mask, st32, steps, np.float64)
    _, eu = rel_l2_pu(got, orc32)
    ep64, _ = rel_l2_pu(got, orc64)
    ep_orc, _ = rel_l2_pu(orc32, orc64)
    assert eu <= 1e-5, (what, eu)
    assert ep64 <= max(1e-5, 2 * ep_orc), (what, ep64, ep_orc)


def test_uneven_pipes_batch(orc):
    """The 64 pipes (5 GB of state) step in cohorts that fit the last-level
    cache.  Two calls: the first (33 steps: two 16-step graph replays and
    one direct launch per cohort) builds the cohorts' tables and graphs, the
    second (34 steps) reuses them from the other buffer parity."""
    steps = 67
    cases = W.uneven_pipes(64, 4, steps)
    assert len(cases) == 64
    with Batch("f32", 2, 64 << 30) as b:
        b.use_graph(True)
        slots = []
        for c in cases:
            prm, mask, st = c.build()
            slots.append(b.add(prm, [a.astype(np.float32) for a in st], mask))
        b.step(33)
        b.step(34)
        b.sync()
        smallest = sorted(range(64), key=lambda k: cases[k].cells)[:3]
        for k, (c, sl) in enumerate(zip(cases, slots)):
            prm, mask, st = c.build()
            st32 = [a.astype(np.float32) for a in st]
            got = b.get_state(sl)
            ref = solo(c, prm, mask, st32, steps, "f32")
            for g, e in zip(got, ref):
                assert np.array_equal(g, e), (k, c.shape)
            if k in smallest:
                check_f32(orc, got, prm, mask, st32, steps, (k, c.shape))


@pytest.mark.parametrize("policy", ["share", "shortest"])
def test_serving_stream(orc, policy):
    """All 32 requests at once: the resident state exceeds the last-level
    cache, so each pump's batches step in cohorts; under the shortest-work
    policy requests are also held and resumed between pumps."""
    cases = W.serving_requests(32, 5, 1.0)
    for c in cases:  # step budgets cut 100-fold: 20 to 200 steps
        c.steps = max(2, c.steps // 100)
    built = {}
    with Server(d1=2, hbm_budget=32 << 30, quantum=64) as srv:
        srv.set_policy(policy, 0)
        owner = {}
        for k, c in enumerate(cases):
            prm, mask, st = c.build()
            owner[srv.submit(prm, st, mask, c.precision, c.steps)] = k
        results = {}
        for _ in range(100000):
            q = srv.query()
            if q["queued"] == 0 and q["running"] == 0:
                break
            srv.pump()
            for comp in srv.completions():
                k = owner[comp.id]
                assert comp.ok and comp.steps == cases[k].steps, k
                results[k] = srv.fetch(comp.id)
                srv.drop(comp.id)
        for comp in srv.completions():
            k = owner[comp.id]
            results[k] = srv.fetch(comp.id)
    assert sorted(results) == list(range(32))
    for k, c in enumerate(cases):
        prm, mask, st = c.build()
        ref = solo(c, prm, mask, st, c.steps, c.precision)
        for g, e in zip(results[k], ref):
            assert np.array_equal(g, e), (k, c.kind, c.shape, c.precision)
        built[k] = (prm, mask, st)
    by_size = sorted(range(32), key=lambda k: cases[k].cells * cases[k].steps)
    k32 = next(k for k in by_size if cases[k].precision == "f32")
    prm, mask, st = built[k32]
    check_f32(orc, results[k32], prm, mask, [a.astype(np.float32) for a in st],
              cases[k32].steps, (k32, cases[k32].shape))
    k16 = next(k for k in by_size if cases[k].precision == "f16s")
    prm, mask, st = built[k16]
    ref64 = oracle_run(orc, prm, mask, st, cases[k16].steps, np.float64)
    ep, eu = rel_l2_pu(results[k16], ref64)
    assert eu <= 4e-3, (k16, cases[k16].shape, eu)
