"""Benchmark of the FSLBM time step on one MI355X.

Workload (default): the turbulent-pipe configuration, 256x256x1024 lattice,
circular wall of radius 120 centred in the cross-section, z periodic with
the pressure jump pin, FP32, Poiseuille + 5% seeded noise initial state
(synthetic; bulk speed 0.05 lattice units, nu = 0.05*2R/5300 for bulk
Re ~ 5300, Blasius-estimated pin).
One step = one FSLBM time step over all 67,108,864 lattice cells (the
kernel reads and writes the 61,341,696 cells of the tiles that hold
fluid; tiles of wall cells hold zeros and are skipped).

Printed: one JSON line with MLUPS (all lattice cells, the paper's
convention), the HBM roofline of the step kernel (against the 8 TB/s peak,
and against the measured rate of a plain device copy) and the CPU baseline
(the C oracle with OpenMP on a bounded sample of the same workload), with
sub-objects for the 16-bit storage modes on the same pipe, the small
32x32x64 FP64 pipe (GPU against the CPU oracle at steps 100 and 1000), the
256^3 Taylor-Green box (FP32 / FP16 / BF16 storage against FP64), the 64
uneven pipes in one batch, and the serving stream.
Other workloads: --workload pipe|tgv|batch|serve|stats.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision f32]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8.0 TB/s
BYTES_PER_VALUE = {"f64": 8, "f32": 4, "f16s": 2, "bf16s": 2}


def algorithmic_bytes_per_cell(precision: str, masked: bool) -> int:
    """B = 2*S + M: the state (4 fields) read once and written once, plus
    the 1-byte mask for a walled domain."""
    return 2 * 4 * BYTES_PER_VALUE[precision] + (1 if masked else 0)


def build_pipe(nx, ny, nz, R, nu, re, noise, seed):
    from flowcontrast_amd import inputs
    mask, nf = inputs.circle_mask(nx, ny, nz, R)
    pin = inputs.blasius_pin(nu, nz, re, R)
    ubar = re * nu / (2 * R)
    st = inputs.poiseuille_state(mask, R, ubar, pin, noise, seed)
    return mask, st, pin, nf


def cpu_baseline(mask, st, pin, nu, steps, precision):
    """Time the C oracle (OpenMP) on `steps` steps of the same workload."""
    from oracle import oracle as O
    dtype = np.float64 if precision == "f64" else np.float32
    prm = O.params(1.0, pin, nu, 0.0, 2)
    sec, threads = O.time_folded(st, prm, mask=mask, nsteps=steps, dtype=dtype)
    n = mask.size
    return {"value": round(n * steps / sec / 1e6, 3), "unit": "MLUPS",
            "cores": threads, "kind": "port",
            "sample": f"{steps} steps of the full {mask.shape[2]}x{mask.shape[1]}x"
                      f"{mask.shape[0]} pipe ({'FP64' if dtype == np.float64 else 'FP32'}),"
                      f" {sec:.1f} s"}


def load_traffic(workload: str, precision: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, scaled by a copy kernel of known size); null if the
    summary for this workload is not present."""
    f = ROOT / "profiles" / "traffic.json"
    if not f.exists():
        return None
    try:
        d = json.loads(f.read_text())
        return d.get(f"{workload}:{precision}", {}).get("bytes_per_launch")
    except Exception:
        return None


def copy_calibration() -> int:
    """A plain copy kernel of known size: 10 copies of 2^30 bytes.  Its
    FETCH_SIZE / WRITE_SIZE counter readings per dispatch divided by 2^30
    give the scale applied to the step kernel's readings."""
    import torch
    n = (1 << 30) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(10):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_dispatch": 1 << 30, "dispatches": 10}))
    return 0


def copy_ceiling() -> float:
    """Read+write rate of a plain device copy of 2^30 bytes (torch's copy
    kernel, 20 copies timed with HIP events on torch's stream): the
    bandwidth a streaming kernel reaches on this GPU, beside the 8 TB/s
    peak."""
    import torch
    n = (1 << 30) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) * 1e-3 / 20
    del a, b
    torch.cuda.empty_cache()
    return 2.0 * (1 << 30) / sec / 1e9


def copy_ceiling_child():
    """copy_ceiling() in a child process (GB/s, or None if it failed)."""
    import subprocess
    try:
        r = subprocess.run([sys.executable, str(Path(__file__).resolve()), "--copy-rate"],
                           capture_output=True, text=True, timeout=300)
        return float(json.loads(r.stdout.strip().splitlines()[-1])["copy_gbs"])
    except Exception as e:  # reported, never fatal to the benchmark
        print(f"bench: copy rate not measured ({type(e).__name__})", file=sys.stderr)
        return None


def rel_l2(a, b) -> float:
    den = float(np.sqrt(sum(np.sum(np.square(x)) for x in b)))
    num = float(np.sqrt(sum(np.sum(np.square(x - y)) for x, y in zip(a, b))))
    return num / den if den > 0 else num


def field_errors(got, ref):
    """Pointwise: relative L2 of the velocity vector field and of the
    pressure deviation from its mean (p is a large constant plus O(u^2)).
    Integral: kinetic energy sum |u|^2 and mean axial velocity.  Past the
    predictability time of a transitional flow the pointwise figures
    measure chaos, the integral ones the solver."""
    gp, rp = got[0] - got[0].mean(), ref[0] - ref[0].mean()
    ke_g = sum(float(np.sum(np.square(a))) for a in got[1:])
    ke_r = sum(float(np.sum(np.square(a))) for a in ref[1:])
    uz_g, uz_r = float(np.mean(got[3])), float(np.mean(ref[3]))
    # the mean axial velocity of a Taylor-Green box is zero: no relative error
    rms = float(np.sqrt(ke_r / (3 * got[3].size)))
    return {"u_rel_l2": rel_l2(got[1:], ref[1:]), "p_rel_l2": rel_l2([gp], [rp]),
            "ke_rel": abs(ke_g - ke_r) / ke_r if ke_r else abs(ke_g),
            "mean_uz_rel": (abs(uz_g - uz_r) / abs(uz_r)
                            if abs(uz_r) > 1e-3 * rms else None)}


def bench_batch(count: int, steps: int, sequential: bool, graph: bool = True):
    """Independent uneven pipes in one batch (one launch per step over a
    table of (pipe, tile, z range) items, autotuned longest-item cap,
    hipGraph replay, or direct launches with graph False) against the same
    pipes solved one after another by
    the single-domain solver.  Timed: the steps only, states resident on
    the device on both sides."""
    from flowcontrast_amd import KERNEL_AUTO, KERNEL_PLAIN, Batch, Solver
    from flowcontrast_amd import workloads as W
    cases = W.uneven_pipes(count, 4, steps)
    cells = sum(c.cells for c in cases)
    print(f"bench: batch of {count} pipes, {cells} cells, {steps} steps", file=sys.stderr,
          flush=True)
    t = time.perf_counter()
    with Batch("f32", 2, 256 << 30) as b:
        slots = []
        for c in cases:
            prm, mask, st = c.build()
            slots.append(b.add(prm, st, mask))
        upload = time.perf_counter() - t
        b.use_graph(graph)
        chunk, _ = b.autotune(10)
        b.sync()
        t0 = time.perf_counter()
        b.step(steps)
        b.sync()
        wall = time.perf_counter() - t0
        finite = all(np.isfinite(a).all() for sl in slots[:4] for a in b.get_state(sl))
    out = {"workload": f"{count} periodic pipes, radius U[12,48], length U[4,10] D, "
                       f"Re U[200,3000], Poiseuille + 2% noise, seed 4, FP32",
           "pipes": count, "cells": cells, "steps_per_solve": steps,
           "solves_per_s": round(count / wall, 4),
           "mlups": round(cells * steps / wall / 1e6, 1),
           "seconds": round(wall, 3), "chunk": chunk, "graph": graph,
           "upload_s": round(upload, 2), "finite": bool(finite)}
    if sequential:
        # the same pipes one after another through the single-domain
        # drop-in: its default (tiled) kernel for all steps, and the plain
        # one-thread-per-cell kernel timed on plain_steps steps per pipe
        # (its per-step cost is constant; solves/s scaled to full solves)
        print("bench: the same pipes one after another", file=sys.stderr, flush=True)
        plain_steps = min(steps, 500)
        seq = seq_plain = 0.0
        for c in cases:
            prm, mask, st = c.build()
            for kern, n in ((KERNEL_AUTO, steps), (KERNEL_PLAIN, plain_steps)):
                with Solver(c.shape, prm, mask=mask, precision="f32", kernel=kern) as s:
                    s.set_state(*st)
                    s.sync()
                    t0 = time.perf_counter()
                    s.step(n)
                    s.sync()
                    dt = time.perf_counter() - t0
                if kern == KERNEL_AUTO:
                    seq += dt
                else:
                    seq_plain += dt * steps / n
        out.update({"sequential_solves_per_s": round(count / seq, 4),
                    "sequential_mlups": round(cells * steps / seq / 1e6, 1),
                    "sequential_seconds": round(seq, 3),
                    "speedup_vs_sequential": round(seq / wall, 3),
                    "sequential_plain_solves_per_s": round(count / seq_plain, 4),
                    "sequential_plain_steps_timed": plain_steps,
                    "speedup_vs_sequential_plain": round(seq_plain / wall, 3)})
    return out


def bench_serve(count: int, rate: float, quantum: int, budget: int, spot: int,
                sequential: bool, policy: str = "share", fill: int = 0):
    """Poisson stream of pipe / Taylor-Green solve requests through the
    serving loop (continuous batching under an HBM budget).  Latency is
    time to solution from each request's scheduled arrival to its
    hand-back.  Then the same requests submitted at once (capacity) and
    one after another through the single-domain solver; the smallest
    requests are checked against solo runs (bitwise) and FP64 runs."""
    from flowcontrast_amd import KERNEL_AUTO, KERNEL_PLAIN, Server, Solver
    from flowcontrast_amd import workloads as W
    cases = W.serving_requests(count, 5, rate)
    built = [c.build() for c in cases]
    work = sum(c.cells * c.steps for c in cases)
    # spot checks: the smallest requests of each precision mode
    spot_ids = set()
    for prec in ("f32", "f16s"):
        ks = [k for k in range(count) if cases[k].precision == prec]
        spot_ids |= set(sorted(ks, key=lambda k: cases[k].cells * cases[k].steps)[:spot])
    print(f"bench: serving {count} requests at {rate}/s", file=sys.stderr, flush=True)

    def run(open_loop: bool):
        lat, status, kept = {}, {}, {}
        peak_used = peak_queue = 0
        with Server(d1=2, hbm_budget=budget, quantum=quantum) as srv:
            srv.set_policy(policy, fill)
            t_start = srv.clock()
            owner, i = {}, 0
            while len(lat) < count:
                now = srv.clock() - t_start
                while i < count and (not open_loop or cases[i].arrival <= now):
                    prm, mask, st = built[i]
                    owner[srv.submit(prm, st, mask, cases[i].precision, cases[i].steps)] = i
                    i += 1
                q = srv.query()
                peak_queue = max(peak_queue, q["queued"])
                if q["queued"] == 0 and q["running"] == 0 and i < count:
                    time.sleep(max(0.0, cases[i].arrival - now))
                    continue
                srv.pump()
                peak_used = max(peak_used, srv.query()["used"])
                for c in srv.completions():
                    k = owner[c.id]
                    t_arr = cases[k].arrival if open_loop else 0.0
                    lat[k] = c.t_done - t_start - t_arr
                    status[k] = c.status
                    if k in spot_ids and c.ok:
                        kept[k] = srv.fetch(c.id)
                    srv.drop(c.id)
            end = srv.clock() - t_start
        return lat, status, kept, end, peak_used, peak_queue

    lat, status, kept, end, peak_used, peak_queue = run(True)
    l = np.array([lat[k] for k in range(count)])
    first = cases[0].arrival
    out = {"workload": f"{count} requests, Poisson {rate}/s, 80% pipes (R U[16,64], "
                       f"L U[4,8] D, Re U[200,6000]) / 20% Taylor-Green (64^3-192^3, "
                       f"Re U[100,1600]), 2000-20000 steps, FP32 or FP16 storage, seed 5",
           "requests": count, "rate_per_s": rate, "quantum": quantum, "policy": policy,
           "fill_cells": fill,
           "hbm_budget": budget, "cell_updates": work,
           "p50_s": round(float(np.percentile(l, 50)), 4),
           "p99_s": round(float(np.percentile(l, 99)), 4),
           "max_s": round(float(l.max()), 4),
           "solves_per_s": round(count / (end - first), 4),
           "failed": int(sum(1 for v in status.values() if v != 0)),
           "peak_hbm_bytes": int(peak_used), "peak_queue": int(peak_queue)}
    _, status2, _, end2, _, _ = run(False)
    out["capacity_solves_per_s"] = round(count / end2, 4)
    out["capacity_mlups"] = round(work / end2 / 1e6, 1)
    checks = []
    for k in sorted(kept):
        c = cases[k]
        prm, mask, st = built[k]
        res = {"request": k, "kind": c.kind, "shape": list(c.shape), "steps": c.steps,
               "precision": c.precision}
        with Solver(c.shape, prm, mask=mask, precision=c.precision, kernel=KERNEL_AUTO) as s:
            s.set_state(*st)
            s.step(c.steps)
            solo = s.get_state()
        res["equal_to_solo"] = bool(all(np.array_equal(a, b) for a, b in zip(kept[k], solo)))
        with Solver(c.shape, prm, mask=mask, precision="f64", kernel=KERNEL_AUTO) as s:
            s.set_state(*st)
            s.step(c.steps)
            ref = s.get_state()
        res.update({k2: (None if v is None else float(f"{v:.3e}"))
                    for k2, v in field_errors(kept[k], ref).items()})
        checks.append(res)
    out["spot_checks_vs_fp64"] = checks
    if sequential:
        # the same requests one after another through the single-domain
        # drop-in, each timed like a served one: state upload, steps and
        # result download (the served capacity run includes all three)
        # The plain one-thread-per-cell kernel is timed on up to 300 steps per
        # request and its step time scaled to the request's budget.
        seq = seq_steps = seq_plain = 0.0
        for c, (prm, mask, st) in zip(cases, built):
            for kern in (KERNEL_AUTO, KERNEL_PLAIN):
                n = c.steps if kern == KERNEL_AUTO else min(c.steps, 300)
                t0 = time.perf_counter()
                with Solver(c.shape, prm, mask=mask, precision=c.precision, kernel=kern) as s:
                    s.set_state(*st)
                    s.sync()
                    t1 = time.perf_counter()
                    s.step(n)
                    s.sync()
                    dt_steps = time.perf_counter() - t1
                    s.get_state()
                dt = time.perf_counter() - t0
                if kern == KERNEL_AUTO:
                    seq += dt
                    seq_steps += dt_steps
                else:
                    seq_plain += dt - dt_steps + dt_steps * c.steps / n
        out["sequential_solves_per_s"] = round(count / seq, 4)
        out["sequential_steps_only_solves_per_s"] = round(count / seq_steps, 4)
        out["capacity_vs_sequential"] = round(seq / end2, 3)
        out["sequential_plain_solves_per_s"] = round(count / seq_plain, 4)
        out["capacity_vs_sequential_plain"] = round(seq_plain / end2, 3)
    return out


def radial_stats(mom, samples, nz, mask2d, R):
    """Radial profiles (unit-width bins of r around the pipe axis) of the
    mean axial velocity and the Reynolds stresses u'u', v'v', w'w' and the
    radial-axial shear u_r'w', from per-(x,y) moment sums over z and
    samples (p, u, v, w, pp, uu, vv, ww, uv, uw, vw)."""
    ny, nx = mask2d.shape
    n = float(samples * nz)
    m = mom / n
    mu, mv, mw = m[1], m[2], m[3]
    uu = m[5] - mu * mu
    vv = m[6] - mv * mv
    ww = m[7] - mw * mw
    uv = m[8] - mu * mv
    uw = m[9] - mu * mw
    vw = m[10] - mv * mw
    yy, xx = np.mgrid[0:ny, 0:nx]
    dx, dy = xx - (nx - 1) / 2.0, yy - (ny - 1) / 2.0
    r = np.hypot(dx, dy)
    with np.errstate(invalid="ignore", divide="ignore"):
        cx, cy = np.where(r > 0, dx / r, 0.0), np.where(r > 0, dy / r, 0.0)
    urw = cx * uw + cy * vw
    fluid = mask2d == 0
    bins = np.floor(r).astype(int)
    nb = int(np.floor(R)) + 1
    out = {}
    for name, f in (("w", mw), ("uu", uu), ("vv", vv), ("ww", ww), ("urw", urw)):
        prof = np.zeros(nb)
        for b_ in range(nb):
            sel = fluid & (bins == b_)
            prof[b_] = f[sel].mean() if sel.any() else 0.0
        out[name] = prof
    return out


def tripped_state(mask, st, R, ubar, amp, seed, noise=0.3, block=16):
    """The Poiseuille state plus a finite-amplitude trip: divergence-free
    streamwise vortices (u, v) = (d psi/dy, -d psi/dx) from a stream
    function psi = (1 - r^2/R^2)^2 cos(m theta + 2 pi k z / nz + phase)
    summed over a few (m, k) modes, scaled to a peak of amp * ubar, a weak
    axial streak (cos theta, one axial wavelength), and seeded random
    velocities of amplitude noise * ubar, constant on blocks of `block`^3
    cells.  On a 128^2 x 512 pipe at bulk Re 5300 the smooth vortices alone
    decayed (the flow stayed laminar and accelerated); with the block noise
    at 30% of the bulk speed the flow became turbulent within 20,000 steps
    and the bulk speed settled at the turbulent value."""
    nz, ny, nx = mask.shape
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:ny, 0:nx].astype(np.float32)
    dx, dy = xx - (nx - 1) / 2, yy - (ny - 1) / 2
    env = np.clip(1 - (dx * dx + dy * dy) / (R * R), 0, None) ** 2
    th = np.arctan2(dy, dx)
    zph = (2 * np.pi / nz) * np.arange(nz, dtype=np.float32)
    u = np.zeros(mask.shape, np.float32)
    v = np.zeros(mask.shape, np.float32)
    for m, k in ((1, 1), (2, 1), (3, 2), (2, 3), (4, 2)):
        ph = np.float32(rng.uniform(0, 2 * np.pi))
        for z in range(nz):
            psi = env * np.cos(m * th + k * zph[z] + ph)
            gy, gx = np.gradient(psi)
            u[z] += gy
            v[z] -= gx
    fl = mask == 0
    peak = float(np.sqrt(u[fl] ** 2 + v[fl] ** 2).max())
    s = amp * ubar / peak
    out = [np.array(a, dtype=np.float64) for a in st]
    out[1] += np.where(fl, s * u, 0.0)
    out[2] += np.where(fl, s * v, 0.0)
    streak = 0.1 * ubar * env * np.cos(th)
    for z in range(nz):
        out[3][z] += np.where(fl[z], streak * np.sin(zph[z]), 0.0)
    if noise > 0:
        nb = [(n + block - 1) // block for n in mask.shape]
        for f in (1, 2, 3):
            c = rng.uniform(-1, 1, nb) * (noise * ubar)
            big = np.repeat(np.repeat(np.repeat(c, block, 0), block, 1), block, 2)
            out[f] += np.where(fl, big[:nz, :ny, :nx], 0.0)
    return out


def bench_stats(spinup: int, window: int, every: int, precisions):
    """The turbulent pipe (256x256x1024, R 120, bulk Re 5300): a tripped
    state (tripped_state) is run in FP32 for `spinup` steps into developed
    turbulence; from that one state every precision mode then runs `window`
    steps, sampling the moments every `every` steps.  Radial profiles of the
    mean axial velocity and the Reynolds stresses are compared with the
    FP64 run's.  Starting all modes from the same turbulent state keeps
    their trajectories together over the window, so the differences are
    those of the arithmetic and storage, not of sampling."""
    from flowcontrast_amd import KERNEL_AUTO, RunParams, Solver
    nx = ny = 256
    nz, R, re, ubar = 1024, 120.0, 5300.0, 0.05
    nu = ubar * 2 * R / re
    mask, st, pin, nf = build_pipe(nx, ny, nz, R, nu, re, 0.05, 3)
    prm = RunParams(tblend=1.0, pin=pin, nu=nu, d1=2, delta=1, blend=0.0)
    fl = mask[0] == 0
    t0 = time.perf_counter()
    st = tripped_state(mask, st, R, ubar, 0.3, 7)
    print(f"bench: stats trip built ({time.perf_counter() - t0:.0f} s)", file=sys.stderr,
          flush=True)
    history = []
    with Solver(mask.shape, prm, mask=mask, precision="f32", kernel=KERNEL_AUTO) as s:
        s.set_state(*st)
        done = 0
        while done < spinup:
            n = min(10000, spinup - done)
            s.step(n)
            done += n
            # bulk velocity and cross-stream rms over the fluid cells
            gs = s.get_state(np.float32)
            ub = -float(np.mean(gs[3][:, fl]))
            cr = float(np.sqrt(np.mean(gs[1][:, fl] ** 2 + gs[2][:, fl] ** 2)))
            history.append((done, round(ub, 5), float(f"{cr:.3e}")))
            print(f"bench: stats spin-up step {done}: bulk u {ub:.5f}, cross-stream rms "
                  f"{cr:.3e} ({time.perf_counter() - t0:.0f} s)", file=sys.stderr, flush=True)
        snap = [np.asarray(a, dtype=np.float64) for a in gs]
    cross = float(np.sqrt(np.mean((snap[1][:, fl] ** 2 + snap[2][:, fl] ** 2))))
    prof = {}
    for prec in precisions:
        with Solver(mask.shape, prm, mask=mask, precision=prec, kernel=KERNEL_AUTO) as s:
            s.set_state(*snap)
            s.moments_reset()
            done = 0
            while done < window:
                n = min(every, window - done)
                s.step(n)
                done += n
                s.moments_accumulate()
            mom, samples = s.moments_read()
            red = s.reduce()
        print(f"bench: stats window {prec} done ({time.perf_counter() - t0:.0f} s)",
              file=sys.stderr, flush=True)
        prof[prec] = radial_stats(mom, samples, nz, mask[0], R)
        prof[prec]["finite"] = bool(red[7] == 0)
    ref = prof["f64"]
    out = {"workload": f"pipe{nx}x{ny}x{nz}_r{int(R)}, Re {re:g}, tripped (vortices 30% of "
                       f"ubar, block noise 30%), FP32 spin-up {spinup} steps, then {window} "
                       f"steps per mode from "
                       f"that state, moments every {every}",
           "spinup": spinup, "window": window, "every": every,
           "spinup_bulk_velocity_cross_rms": history, "cross_stream_rms_at_start": round(cross, 6),
           "errors_vs_f64": {}}
    wmax = float(np.abs(ref["w"]).max())
    for prec in precisions:
        if prec == "f64":
            continue
        e = {}
        for name in ("w", "uu", "vv", "ww", "urw"):
            a, b_ = prof[prec][name], ref[name]
            scale = wmax if name == "w" else float(np.abs(b_).max())
            e[name] = float(f"{float(np.abs(a - b_).max()) / scale:.3e}") if scale > 0 else None
        e["finite"] = prof[prec]["finite"]
        out["errors_vs_f64"][prec] = e
    out["f64_profile_w_axis_wall"] = [float(ref["w"][0]), float(ref["w"][-1])]
    out["f64_max_stress"] = {k: float(np.abs(ref[k]).max()) for k in ("uu", "vv", "ww", "urw")}
    out["f64_profiles"] = {k: [float(f"{v:.5e}") for v in ref[k]] for k in ("w", "uu", "vv",
                                                                          "ww", "urw")}
    return out


def bench_small_pipe():
    """The small-pipe case: 32x32x64, radius 14, FP64, bulk Re 200 (nu 0.01),
    Poiseuille + 1% noise; 1000 steps on the GPU (tiled kernel) and in the
    C oracle on the host cores, compared at steps 100 and 1000.  The GPU
    rate is timed over 1000 steps after 200 untimed ones."""
    from flowcontrast_amd import KERNEL_AUTO, RunParams, Solver, inputs
    from oracle import oracle as O
    nx = ny = 32
    nz, R, nu = 64, 14.0, 0.01
    ubar = 200 * nu / (2 * R)
    pin = inputs.laminar_pin(nu, nz, ubar, R)
    mask, _ = inputs.circle_mask(nx, ny, nz, R)
    st = inputs.poiseuille_state(mask, R, ubar, pin, 0.01, 1)
    prm = RunParams(tblend=1.0, pin=pin, nu=nu, d1=2, delta=1, blend=0.0)
    oprm = O.params(1.0, pin, nu, 0.0, 2)
    cells = nx * ny * nz
    out = {"workload": "pipe32x32x64_r14, FP64, Re 200, 1000 steps", "cells": cells}
    with Solver(mask.shape, prm, mask=mask, precision="f64", kernel=KERNEL_AUTO) as s:
        s.set_state(*st)
        s.step(100)
        g100 = s.get_state()
        s.step(900)
        g1000 = s.get_state()
        s.set_state(*st)
        s.step(200)  # untimed warm-up: the clock ramps up over the first launches
        s.sync()
        ms = s.time_steps(1000)
    out["gpu_mlups"] = round(cells * 1000 / ms / 1e3, 1)
    t0 = time.perf_counter()
    r100 = O.step_folded(st, oprm, mask=mask, nsteps=100)
    r1000 = O.step_folded(r100, oprm, mask=mask, nsteps=900)
    cpu_s = time.perf_counter() - t0
    out["cpu_mlups"] = round(cells * 1000 / cpu_s / 1e6, 1)
    out["cpu_threads"] = O.max_threads()

    def maxrel(g, r):
        return max(float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)) for a, b in zip(g, r))

    out["max_rel_diff_step100"] = float(f"{maxrel(g100, r100):.3e}")
    out["max_rel_diff_step1000"] = float(f"{maxrel(g1000, r1000):.3e}")
    return out


def bench_rough(steps: int, warmup: int, precisions=("f32", "f16s", "bf16s"),
                plain_steps: int = 20, chunk: int = -1):
    """A rough-walled pipe (a mask that varies along z) of the headline
    pipe's size: 256x256x1024, mean radius 120, wall corrugation of 4 cells
    and 600 seeded bumps, bulk Re 5300; Poiseuille + 5% noise.  Step rate of
    the tiled kernel (its variant for such masks) in each 32-bit mode and of
    the plain kernel (FP32), over at least 200 steps after at least 50
    warm-up steps (the plain kernel over `plain_steps`, none if 0).  The
    tiled schedule is autotuned per mode unless `chunk` > 0 fixes it."""
    from flowcontrast_amd import KERNEL_PLAIN, KERNEL_TILED, RunParams, Solver, inputs
    nx = ny = 256
    nz, R, amp = 1024, 120.0, 4.0
    re, ubar = 5300.0, 0.05
    nu = ubar * 2 * R / re
    pin = inputs.blasius_pin(nu, nz, re, R)
    mask, nf = inputs.rough_pipe_mask(nx, ny, nz, R, amp, 600, 5)
    st = [a.astype(np.float32) for a in inputs.poiseuille_state(mask, R, ubar, pin, 0.05, 3)]
    prm = RunParams(tblend=1.0, pin=pin, nu=nu, d1=2, delta=1, blend=0.0)
    n = mask.size
    steps, warmup = max(steps, 200), max(warmup, 50)
    out = {"workload": f"rough_pipe{nx}x{ny}x{nz}_r{int(R)}_amp{int(amp)}_600bumps",
           "fluid_cells": nf, "steps": steps, "warmup": warmup,
           "bytes_per_cell_read_by_kernel": "state + 2-byte wall code"}
    for prec in precisions:
        with Solver(mask.shape, prm, mask=mask, precision=prec) as s:
            assert s.kernel == KERNEL_TILED
            s.set_state(*st)
            if chunk > 0:
                s.set_tile_chunk(chunk)
            elif chunk < 0:
                s.autotune(50)
            s.step(warmup)
            s.sync()
            t0 = time.perf_counter()
            ms = s.time_steps(steps) / steps
            s.sync()
            wall = time.perf_counter() - t0
            if s.reduce()[7] != 0:
                raise RuntimeError(f"non-finite values in the {prec} rough-pipe state")
            B = algorithmic_bytes_per_cell(prec, True)
            mlups = n * steps / wall / 1e6
            out[prec] = {"mlups": round(mlups, 1), "kernel_ms": round(ms, 4),
                         "chunk": s.tile_chunk, "cells_per_launch": s.step_cells,
                         "roofline_frac": round(B * s.step_cells / (ms * 1e-3) / 1e9
                                                / HBM_PEAK_GBS, 4)}
            if prec != "f32" and "f32" in out:
                out[prec]["vs_f32"] = round(mlups / out["f32"]["mlups"], 3)
    if plain_steps <= 0 or "f32" not in out:
        return out
    with Solver(mask.shape, prm, mask=mask, precision="f32", kernel=KERNEL_PLAIN) as s:
        s.set_state(*st)
        s.step(2)
        s.sync()
        t0 = time.perf_counter()
        ms = s.time_steps(plain_steps) / plain_steps
        s.sync()
        wall = time.perf_counter() - t0
        out["plain_f32"] = {"mlups": round(n * plain_steps / wall / 1e6, 1),
                            "kernel_ms": round(ms, 4), "steps": plain_steps}
    out["tiled_vs_plain_f32"] = round(out["f32"]["mlups"] / out["plain_f32"]["mlups"], 2)
    return out


def bench_storage_modes(mask, st, prm, steps: int, warmup: int):
    """The 16-bit storage modes (FP32 arithmetic) and FP64 (the parity
    path) on the same pipe and initial state as the FP32 line: step rate,
    roofline fraction over their bytes per cell (17, 65), and the ratio to
    an FP32 run timed the same way.  Each mode is timed over at least 200 steps after at least 50
    warm-up steps (the ratio of rates over a few steps is mostly noise)."""
    from flowcontrast_amd import KERNEL_AUTO, Solver
    n = mask.size
    steps, warmup = max(steps, 200), max(warmup, 50)
    out = {"workload": "the FP32 line's pipe and initial state", "steps": steps,
           "warmup": warmup}
    for prec in ("f32", "f16s", "bf16s", "f64"):
        with Solver(mask.shape, prm, mask=mask, precision=prec, kernel=KERNEL_AUTO) as s:
            s.set_state(*(st if prec == "f64" else [a.astype(np.float32) for a in st]))
            s.autotune(50)  # on the uploaded state, right before the warm-up
            s.step(warmup)
            s.sync()
            t0 = time.perf_counter()
            ms = s.time_steps(steps) / steps
            s.sync()
            wall = time.perf_counter() - t0
            B = algorithmic_bytes_per_cell(prec, True)
            mlups = n * steps / wall / 1e6
            out[prec] = {"mlups": round(mlups, 1), "kernel_ms": round(ms, 4),
                         "chunk": s.tile_chunk,
                         "roofline_frac": round(B * s.step_cells / (ms * 1e-3) / 1e9
                                                / HBM_PEAK_GBS, 4),
                         "vs_f32": round(mlups / out["f32"]["mlups"], 3) if prec != "f32"
                         else 1.0}
    return out


def bench_tgv(n: int, steps: int, every: int, timed_steps: int, precisions):
    """Taylor-Green vortex on an n^3 periodic box (U0 = 0.05, Re 1600):
    step rate and roofline of the unmasked kernel, and the kinetic-energy
    decay KE(t) and dissipation rate -dKE/dt of each precision against the
    FP64 run."""
    from flowcontrast_amd import KERNEL_AUTO, RunParams, Solver, inputs
    nu = inputs.tgv_viscosity(n, 0.05, 1600.0)
    prm = RunParams(tblend=1.0, pin=0.0, nu=nu, d1=2, delta=1, blend=0.0)
    st = inputs.taylor_green_state(n, 0.05, 1.0)
    cells = n ** 3
    out = {"workload": f"taylor_green{n}^3, U0 0.05, Re 1600, nu {nu:.4e}, {steps} steps",
           "cells": cells, "steps": steps}
    ke = {}
    for prec in precisions:
        with Solver((n, n, n), prm, mask=None, precision=prec, kernel=KERNEL_AUTO) as s:
            s.set_state(*st)
            series = []
            for t in range(0, steps + 1, every):
                if t:
                    s.step(every)
                r = s.reduce()  # sums of p, u, v, w, u^2+v^2+w^2, ...
                series.append(0.5 * r[5] / cells)
            ke[prec] = np.array(series)
            if prec != "f64":
                s.autotune(10)
                ms = s.time_steps(timed_steps) / timed_steps
                B = algorithmic_bytes_per_cell(prec, False)
                out[f"{prec}_mlups"] = round(cells / ms / 1e3, 1)
                out[f"{prec}_roofline_frac"] = round(B * cells / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    ref = ke["f64"]
    eps_ref = -np.diff(ref) / every
    out["ke0"] = float(ref[0])
    out["ke_final_over_ke0_f64"] = float(ref[-1] / ref[0])
    # the breakdown to turbulence: where the dissipation rate peaks
    k_ref = int(np.argmax(eps_ref))
    out["f64_peak_dissipation_step"] = (k_ref + 1) * every
    out["f64_peak_dissipation"] = float(f"{float(eps_ref[k_ref]):.6e}")
    for prec in precisions:
        if prec == "f64":
            continue
        eps = -np.diff(ke[prec]) / every
        out[f"{prec}_ke_max_rel_err"] = float(f"{float(np.max(np.abs(ke[prec] - ref) / ref)):.3e}")
        out[f"{prec}_dissipation_max_rel_err"] = float(
            f"{float(np.max(np.abs(eps - eps_ref)) / np.max(np.abs(eps_ref))):.3e}")
        k = int(np.argmax(eps))
        out[f"{prec}_peak_dissipation_step"] = (k + 1) * every
        out[f"{prec}_peak_dissipation_rel_err"] = float(
            f"{abs(float(eps[k]) - float(eps_ref[k_ref])) / float(eps_ref[k_ref]):.3e}")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", default="f32", choices=list(BYTES_PER_VALUE))
    ap.add_argument("--kernel", default="auto", choices=["auto", "plain", "tiled"])
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--nz", type=int, default=1024)
    ap.add_argument("--radius", type=float, default=120.0)
    ap.add_argument("--chunk", type=int, default=-1,
                    help="tiled kernel: z planes per work item (0 = library "
                         "default, -1 = autotune)")
    ap.add_argument("--cpu-steps", type=int, default=25,
                    help="steps of the CPU baseline sample (about 10 s on 16 host threads)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="all",
                    choices=["all", "pipe", "batch", "serve", "stats", "tgv", "rough"],
                    help="all: the pipe step line with the batched and serving "
                         "results attached")
    ap.add_argument("--rough-precisions", default="f32,f16s,bf16s",
                    help="--workload rough: storage modes to time")
    ap.add_argument("--rough-plain-steps", type=int, default=20,
                    help="--workload rough: plain-kernel steps timed (0: none)")
    ap.add_argument("--batch-count", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=10000)
    ap.add_argument("--batch-direct", action="store_true",
                    help="--workload batch: direct launches instead of hipGraph replay")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--serve-rate", type=float, default=1.0)
    ap.add_argument("--serve-quantum", type=int, default=200)
    ap.add_argument("--serve-fill", type=int, default=0,
                    help="shortest policy: cells stepped per pump at least (0: library default)")
    ap.add_argument("--serve-policy", default="share", choices=["share", "shortest"],
                    help="which resident requests step in a pump: least work left "
                         "first up to a GPU's fill, or all of them")
    ap.add_argument("--serve-budget", type=int, default=32 << 30)
    ap.add_argument("--spot-checks", type=int, default=2,
                    help="serving: requests of each precision checked against FP64")
    ap.add_argument("--no-sequential", action="store_true")
    ap.add_argument("--tgv-steps", type=int, default=5000)
    ap.add_argument("--stats-spinup", type=int, default=150000,
                    help="stats: FP32 steps from the tripped state into turbulence")
    ap.add_argument("--stats-window", type=int, default=20000)
    ap.add_argument("--stats-every", type=int, default=100)
    ap.add_argument("--stats-precisions", default="f64,f32,f16s,bf16s")
    ap.add_argument("--no-ceiling", action="store_true",
                    help="skip the plain-copy rate beside the roofline (torch's kernels "
                         "crash the host side of rocprofv3 --kernel-trace)")
    ap.add_argument("--copy-rate", action="store_true",
                    help="print only the plain-copy rate (run by the benchmark in a child "
                         "process)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain device copy of 1 GiB x 10 (the "
                         "scale for the FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.copy_rate:
        print(json.dumps({"copy_gbs": copy_ceiling()}), flush=True)
        return 0

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: the FSLBM step runs on one GPU; --gpus {args.gpus} "
              f"(WORLD_SIZE={world}) is not supported. Independent pipes or "
              f"requests shard one per GPU; a single domain would need a "
              f"z-slab halo exchange, which is not implemented.",
              file=sys.stderr)
        return 2

    if args.workload == "batch":
        print(json.dumps(bench_batch(args.batch_count, args.batch_steps,
                                     not args.no_sequential, not args.batch_direct)), flush=True)
        return 0
    if args.workload == "tgv":
        print(json.dumps(bench_tgv(256, args.tgv_steps, 250, 100, ["f64", "f32", "f16s", "bf16s"])),
              flush=True)
        return 0
    if args.workload == "rough":
        print(json.dumps(bench_rough(args.steps, args.warmup, args.rough_precisions.split(","),
                                     args.rough_plain_steps, args.chunk)), flush=True)
        return 0
    if args.workload == "stats":
        print(json.dumps(bench_stats(args.stats_spinup, args.stats_window, args.stats_every,
                                     args.stats_precisions.split(","))), flush=True)
        return 0
    if args.workload == "serve":
        print(json.dumps(bench_serve(args.serve_requests, args.serve_rate, args.serve_quantum,
                                     args.serve_budget, args.spot_checks,
                                     not args.no_sequential, args.serve_policy,
                                     args.serve_fill)), flush=True)
        return 0

    from flowcontrast_amd import KERNEL_AUTO, KERNEL_PLAIN, KERNEL_TILED, RunParams, Solver

    nx = ny = args.nx
    nz = args.nz
    R = args.radius
    re, ubar = 5300.0, 0.05
    nu = ubar * 2 * R / re
    mask, st, pin, nf = build_pipe(nx, ny, nz, R, nu, re, 0.05, 3)
    prm = RunParams(tblend=1.0, pin=pin, nu=nu, d1=2, delta=1, blend=0.0)
    kern = {"auto": KERNEL_AUTO, "plain": KERNEL_PLAIN, "tiled": KERNEL_TILED}[args.kernel]
    n = nx * ny * nz
    workload = f"pipe{nx}x{ny}x{nz}_r{int(R)}"

    # The copy rate runs in a child process before this one touches the
    # GPU: torch carries its own HIP runtime, which finds no device in a
    # process where libfslbm's runtime already holds it.
    ceiling = None if args.no_ceiling else copy_ceiling_child()
    s = Solver(mask.shape, prm, mask=mask, precision=args.precision, kernel=kern)
    if args.precision == "f64":
        s.set_state(*st)
    else:
        s.set_state(*[a.astype(np.float32) for a in st])
    # The schedule is tuned on the uploaded state (the trial launches do not
    # advance it), right before the warm-up: the host-side encode and upload
    # leave the GPU idle for a while, and the tuning launches bring it back
    # to its working clock before the timed region.
    if args.chunk > 0:
        s.set_tile_chunk(args.chunk)
    elif args.chunk < 0:
        s.autotune(50)
    kernel_name = {1: "plain", 2: "tiled"}[s.kernel]
    print(f"bench: {workload} {args.precision} kernel={kernel_name} "
          f"warmup={args.warmup} steps={args.steps}", file=sys.stderr, flush=True)

    s.step(args.warmup)
    s.sync()
    # Timed region: exactly K steps, synchronised on both sides.  The HIP
    # events are recorded on the context's stream, where the step kernel
    # runs (one launch per step).
    t0 = time.perf_counter()
    ev_ms = s.time_steps(args.steps)
    s.sync()
    wall = time.perf_counter() - t0
    red = s.reduce()
    if red[7] != 0:
        print("bench: non-finite values in the state", file=sys.stderr)
        return 1

    ms_per_step = wall * 1e3 / args.steps
    mlups = n * args.steps / wall / 1e6
    kernel_ms = ev_ms / args.steps
    B = algorithmic_bytes_per_cell(args.precision, True)
    # The tiled kernel reads and writes the cells of the tiles that hold
    # fluid; tiles of wall cells are skipped (their state stays zero).
    cells = s.step_cells
    achieved = B * cells / (kernel_ms * 1e-3) / 1e9
    out = {
        "metric": "MLUPS (all lattice cells) of the FSLBM step, 1 GPU, with % HBM roofline",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f64": "f64", "f32": "f32", "f16s": "f16-storage/f32",
                  "bf16s": "bf16-storage/f32"}[args.precision],
        "data": "synthetic (Poiseuille + 5% seeded noise)",
        "config": {"workload": workload, "lattice": [nx, ny, nz], "radius": R,
                   "fluid_cells_per_section": int(nf), "re_bulk": re, "nu": nu,
                   "pin": pin, "d1": 2, "blend": 0.0, "kernel": kernel_name,
                   "chunk": s.tile_chunk,
                   "precision": args.precision},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": load_traffic(workload, args.precision),
                     "bytes_per_cell": B, "cells_per_launch": cells,
                     "kernel_ms": round(kernel_ms, 4)},
        "mflups_fluid": round(nf * nz * args.steps / wall / 1e6, 1),
    }
    s.close()
    if ceiling:
        out["roofline"]["copy_gbs"] = round(ceiling, 1)
        out["roofline"]["frac_of_copy"] = round(achieved / ceiling, 4)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mask, st, pin, nu, args.cpu_steps,
                                           args.precision)
    if args.workload == "all" and args.precision == "f32":
        try:
            out["storage_modes"] = bench_storage_modes(mask, st, prm, args.steps, args.warmup)
        except Exception as e:  # reported in the line, never hidden
            out["storage_modes"] = {"error": f"{type(e).__name__}: {e}"}
    if args.workload == "all":
        for key, fn in (("small_pipe_f64", bench_small_pipe),
                        ("rough_pipe", lambda: bench_rough(args.steps, args.warmup)),
                        ("taylor_green", lambda: bench_tgv(256, 5000, 250, 100,
                                                          ["f64", "f32", "f16s", "bf16s"])),
                        ("batched", lambda: bench_batch(args.batch_count, args.batch_steps,
                                                       not args.no_sequential)),
                        ("serving", lambda: bench_serve(args.serve_requests, args.serve_rate,
                                                        args.serve_quantum, args.serve_budget,
                                                        args.spot_checks,
                                                        not args.no_sequential,
                                                        args.serve_policy,
                                                        args.serve_fill))):
            try:
                out[key] = fn()
            except Exception as e:  # reported in the line, never hidden
                out[key] = {"error": f"{type(e).__name__}: {e}"}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
