"""GPU tests of batched domains: a domain gives bitwise the same state
alone (single context, tiled kernel) and inside a batch, whatever else
the batch holds, with or without hipGraph replay, and when domains are
admitted or removed while others run; admission respects the HBM
budget."""
import numpy as np
import pytest

from flowcontrast_amd import (KERNEL_TILED, Batch, BudgetFull, RunParams, Solver,
                              inputs)

pytestmark = pytest.mark.gpu


def pipe_case(n, nz, R, re, nu, seed, blend=0.0, tb=1.0):
    mask, _ = inputs.circle_mask(n, n, nz, R)
    ubar = re * nu / (2 * R)
    pin = inputs.laminar_pin(nu, nz, ubar, R)
    st = inputs.poiseuille_state(mask, R, ubar, pin, 0.02, seed)
    return RunParams(tblend=tb, pin=pin, nu=nu, d1=2, delta=1, blend=blend), mask, st


def tgv_case(n):
    nu = inputs.tgv_viscosity(n, 0.05, 800.0)
    return RunParams(tblend=1.0, pin=0.0, nu=nu, d1=2, delta=1, blend=0.0), None, \
        inputs.taylor_green_state(n, 0.05, 1.0)


def alone(prm, mask, st, steps, precision):
    shape = np.shape(st[0])
    with Solver(shape, prm, mask=mask, precision=precision, kernel=KERNEL_TILED) as s:
        s.set_state(*st)
        s.step(steps)
        return s.get_state()


CASES = [pipe_case(30, 64, 12.5, 300, 0.01, 1),
         pipe_case(52, 40, 23.0, 900, 0.008, 2, blend=0.5, tb=0.9),
         pipe_case(100, 24, 47.0, 2500, 0.005, 3),
         tgv_case(32)]


@pytest.mark.parametrize("precision", ["f32", "f16s"])
@pytest.mark.parametrize("graph", [False, True])
def test_batch_equals_alone(precision, graph):
    # with graphs: two replays of the 16-step graph and 5 direct launches
    steps = 37 if graph else 13
    ref = [alone(prm, mask, st, steps, precision) for prm, mask, st in CASES]
    with Batch(precision, 2, 8 << 30) as b:
        b.use_graph(graph)
        slots = [b.add(prm, st, mask) for prm, mask, st in CASES]
        b.step(steps)
        for slot, r in zip(slots, ref):
            got = b.get_state(slot)
            for g, e in zip(got, r):
                assert np.array_equal(g, e)


def test_admission_and_removal_mid_run():
    """Domains admitted later or next to removed ones still match a solo
    run of the same number of steps."""
    prmA, mA, sA = CASES[0]
    prmB, mB, sB = CASES[2]
    prmC, mC, sC = CASES[3]
    refA = alone(prmA, mA, sA, 20, "f32")
    refB = alone(prmB, mB, sB, 7, "f32")
    refC = alone(prmC, mC, sC, 11, "f32")
    with Batch("f32", 2, 8 << 30) as b:
        b.use_graph(True)
        a = b.add(prmA, sA, mA)
        b.step(9)
        c = b.add(prmC, sC, mC)
        b.step(4)
        bb = b.add(prmB, sB, mB)
        b.step(7)
        assert all(np.array_equal(g, e) for g, e in zip(b.get_state(c), refC))
        b.remove(c)
        assert all(np.array_equal(g, e) for g, e in zip(b.get_state(bb), refB))
        b.remove(bb)
        assert all(np.array_equal(g, e) for g, e in zip(b.get_state(a), refA))
        assert b.query()["domains"] == 1


def test_hbm_budget_admission():
    prm, mask, st = CASES[2]
    need = Batch.domain_bytes("f32", np.shape(st[0]), True)
    with Batch("f32", 2, int(need * 2.5)) as b:
        s0 = b.add(prm, st, mask)
        s1 = b.add(prm, st, mask)
        with pytest.raises(BudgetFull):
            b.add(prm, st, mask)
        assert b.query()["used"] == 2 * need
        b.remove(s0)
        s2 = b.add(prm, st, mask)  # memory freed: admitted into the free slot
        assert s2 == s0 and b.query()["domains"] == 2
        b.remove(s1)
        b.remove(s2)
        assert b.query()["used"] == 0


def test_batch_autotune_keeps_state():
    with Batch("f32", 2, 8 << 30) as b:
        slots = [b.add(prm, st, mask) for prm, mask, st in CASES]
        before = [b.get_state(s) for s in slots]
        chunk, ms = b.autotune(2)
        assert ms > 0
        for s, bef in zip(slots, before):
            assert all(np.array_equal(g, e) for g, e in zip(b.get_state(s), bef))
        b.step(5)
        for (prm, mask, st), s in zip(CASES, slots):
            ref = alone(prm, mask, st, 5, "f32")
            assert all(np.array_equal(g, e) for g, e in zip(b.get_state(s), ref))


@pytest.mark.parametrize("precision", ["f32", "f16s"])
def test_periodic_partial_tiles_are_deterministic(precision):
    """A periodic box whose sides are not multiples of the tile (88 = 64 +
    24 lanes, 88 rows = 7 x 12 + 4): lanes and rows past the box wrap onto
    cells another workgroup owns and must not write them.  Two solo runs
    and a batched run agree bitwise."""
    prm, mask, st = tgv_case(88)
    a = alone(prm, mask, st, 60, precision)
    b = alone(prm, mask, st, 60, precision)
    assert all(np.array_equal(g, e) for g, e in zip(a, b))
    with Batch(precision, 2, 8 << 30) as bt:
        sl = bt.add(prm, st, mask)
        bt.step(60)
        assert all(np.array_equal(g, e) for g, e in zip(bt.get_state(sl), a))


@pytest.mark.parametrize("precision", ["f32", "bf16s"])
def test_batch_d1_1_equals_alone(precision):
    """A batch of d1 = 1 domains (corrector spacing 1) with both kernel
    classes (blend 0 and blend != 0) matches solo runs bitwise."""
    cases = [pipe_case(30, 40, 12.5, 300, 0.01, 5),
             pipe_case(26, 33, 10.0, 200, 0.01, 6, blend=1.0, tb=0.8)]
    cases = [(RunParams(tblend=p.tblend, pin=p.pin, nu=p.nu, d1=1, delta=1, blend=p.blend), m, s)
             for p, m, s in cases]
    ref = [alone(prm, mask, st, 17, precision) for prm, mask, st in cases]
    with Batch(precision, 1, 8 << 30) as b:
        b.use_graph(True)
        slots = [b.add(prm, st, mask) for prm, mask, st in cases]
        b.step(17)
        for slot, r in zip(slots, ref):
            assert all(np.isfinite(e).all() for e in r)
            assert all(np.array_equal(g, e) for g, e in zip(b.get_state(slot), r))


@pytest.mark.parametrize("precision", ["f64", "f32", "f16s"])
def test_tiny_work_items_change_only_the_schedule(precision):
    """Work items of 1 to 5 z planes (the solo kernel's chunk, the batch's
    longest-item cap) give bitwise the state of the default schedule: a
    cell's arithmetic does not depend on which item holds it."""
    cases = [CASES[1], CASES[3]]  # order-4 arms in a pipe; a periodic box
    refs = [alone(prm, mask, st, 6, precision) for prm, mask, st in cases]
    for (prm, mask, st), ref in zip(cases, refs):
        for ch in (1, 2, 3, 5):
            with Solver(np.shape(st[0]), prm, mask=mask, precision=precision,
                        kernel=KERNEL_TILED) as s:
                s.set_tile_chunk(ch)
                s.set_state(*st)
                s.step(6)
                got = s.get_state()
            assert all(np.array_equal(g, e) for g, e in zip(got, ref)), ch
    if precision == "f64":  # batches take the 32-bit arithmetic modes
        return
    for cap in (1, 3):
        with Batch(precision, 2, 8 << 30) as b:
            slots = [b.add(prm, st, mask) for prm, mask, st in cases]
            b.set_chunk(cap)
            b.step(6)
            for sl, ref in zip(slots, refs):
                assert all(np.array_equal(g, e) for g, e in zip(b.get_state(sl), ref)), cap
