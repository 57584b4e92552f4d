"""GPU tests of the serving loop: every request's result is bitwise the
same as a solo run of the single-domain tiled path in its precision mode,
whatever else is resident; admission follows arrival order under the HBM
budget; a request that cannot run fails alone."""
import numpy as np
import pytest

from flowcontrast_amd import (KERNEL_TILED, FslbmError, RunParams, Server, Solver,
                              inputs)
from flowcontrast_amd.batch import Batch

pytestmark = pytest.mark.gpu


def pipe_req(R, nz, re, seed, ubar=0.05, noise=0.02):
    n = int(2 * R) + 4
    mask, _ = inputs.circle_mask(n, n, nz, R)
    nu = ubar * 2 * R / re
    pin = inputs.laminar_pin(nu, nz, ubar, R)
    st = inputs.poiseuille_state(mask, R, ubar, pin, noise, seed)
    return RunParams(tblend=1.0, pin=pin, nu=nu, d1=2, delta=1, blend=0.0), mask, st


def tgv_req(n, re):
    nu = inputs.tgv_viscosity(n, 0.05, re)
    return RunParams(tblend=1.0, pin=0.0, nu=nu, d1=2, delta=1, blend=0.0), None, \
        inputs.taylor_green_state(n, 0.05, 1.0)


def alone(prm, mask, st, steps, precision):
    with Solver(np.shape(st[0]), prm, mask=mask, precision=precision,
                kernel=KERNEL_TILED) as s:
        s.set_state(*st)
        s.step(steps)
        return s.get_state()


REQS = [(pipe_req(10.0, 80, 400, 11), "f32", 37),
        (tgv_req(32, 800), "f16s", 50),
        (pipe_req(16.5, 48, 1500, 12), "f16s", 23),
        (pipe_req(7.0, 64, 250, 13), "f32", 64),
        (tgv_req(24, 200), "f32", 9)]


def drain(srv, limit=10000):
    done = []
    for _ in range(limit):
        q = srv.query()
        if q["queued"] == 0 and q["running"] == 0:
            break
        srv.pump()
        done += srv.completions()
    done += srv.completions()
    return done


@pytest.mark.parametrize("policy, fill", [("share", 0), ("shortest", 0), ("shortest", 1)])
def test_results_equal_solo_runs(policy, fill):
    """With fill 1 the shortest-work policy steps one request per pump (and
    any that waited too long), so requests are paused and resumed many
    times; results stay bitwise those of solo runs."""
    with Server(d1=2, hbm_budget=8 << 30, quantum=16) as srv:
        srv.set_policy(policy, fill)
        ids = [srv.submit(prm, st, mask, prec, steps)
               for (prm, mask, st), prec, steps in REQS]
        done = drain(srv)
        assert sorted(c.id for c in done) == sorted(ids)
        for c, ((prm, mask, st), prec, steps) in zip(sorted(done, key=lambda c: c.id), REQS):
            assert c.ok and c.steps == steps
            assert c.t_submit <= c.t_admit <= c.t_done
            got = srv.fetch(c.id)
            ref = alone(prm, mask, st, steps, prec)
            for g, e in zip(got, ref):
                assert np.array_equal(g, e), (c.id, prec, steps)
            srv.drop(c.id)
        assert srv.query()["used"] == 0


def test_arrivals_while_running_join_the_batch():
    """A request submitted while others are resident is encoded by the
    server's encoder thread meanwhile and admitted at a pump after that,
    while the first is still running; its result is unchanged by the
    company it ran in."""
    (prmA, mA, sA), _, _ = REQS[0]
    (prmB, mB, sB), _, _ = REQS[2]
    with Server(d1=2, hbm_budget=8 << 30, quantum=8) as srv:
        a = srv.submit(prmA, sA, mA, "f32", 4000)
        srv.pump()  # nothing to step: waits for the encoder, then uploads a
        assert srv.query()["running"] == 1
        srv.pump()
        b = srv.submit(prmB, sB, mB, "f32", 30)
        for _ in range(400):  # a has 498 pumps left
            srv.pump()
            if srv.query()["running"] == 2:
                break
        assert srv.query()["running"] == 2
        done = {c.id: c for c in drain(srv)}
        assert done[a].ok and done[b].ok
        assert done[b].t_admit > done[a].t_admit
        for rid, (prm, mask, st), steps in ((a, (prmA, mA, sA), 4000), (b, (prmB, mB, sB), 30)):
            ref = alone(prm, mask, st, steps, "f32")
            assert all(np.array_equal(g, e) for g, e in zip(srv.fetch(rid), ref))


def test_budget_admits_in_arrival_order():
    (prm, mask, st), _, _ = REQS[2]
    need = Batch.domain_bytes("f32", np.shape(st[0]), True)
    with Server(d1=2, hbm_budget=int(2.5 * need), quantum=10) as srv:
        ids = [srv.submit(prm, st, mask, "f32", 2000) for _ in range(3)]
        for _ in range(150):  # the encoder readies the first two in turn
            srv.pump()
            if srv.query()["running"] == 2:
                break
        q = srv.query()
        assert q["running"] == 2 and q["queued"] == 1 and q["used"] == 2 * need
        done = {c.id: c for c in drain(srv)}
        assert all(done[i].ok for i in ids)
        # the third waited for memory: admitted after the first two finished
        assert done[ids[2]].t_admit >= max(done[ids[0]].t_done, done[ids[1]].t_done)
        r0 = srv.fetch(ids[0])
        for i in ids[1:]:
            assert all(np.array_equal(g, e) for g, e in zip(srv.fetch(i), r0))


def test_invalid_requests():
    (prm, mask, st), _, _ = REQS[0]
    need = Batch.domain_bytes("f32", np.shape(st[0]), True)
    with Server(d1=2, hbm_budget=need, quantum=10) as srv:
        longer = [np.pad(a, ((0, 8), (0, 0), (0, 0)), mode="edge") for a in st]
        with pytest.raises(FslbmError):  # larger than the whole budget
            srv.submit(prm, longer, np.pad(mask, ((0, 8), (0, 0), (0, 0)), mode="edge"),
                       "f32", 10)
        with pytest.raises(FslbmError):  # FP64 is not a serving mode
            srv.submit(prm, st, mask, "f64", 10)
        with pytest.raises(FslbmError):
            srv.submit(prm, st, mask, "f32", 0)
        with pytest.raises(FslbmError):  # unknown policy, negative fill
            srv.set_policy("shortest", -1)
        from flowcontrast_amd._lib import lib
        assert lib().fslbm_server_set_policy(srv._h, 7, 0) < 0
        # too short for the tiled kernels: fails alone, at submission
        short = [a[:8] for a in st]
        bad = srv.submit(prm, short, mask[:8], "f32", 10)
        good = srv.submit(prm, st, mask, "f32", 10)
        done = {c.id: c for c in drain(srv)}
        assert done[bad].status < 0 and done[good].ok
