"""Benchmark of the FSLBM time step on one MI355X.

Workload (default): the turbulent-pipe configuration, 256x256x1024 lattice,
circular wall of radius 120 centred in the cross-section, z periodic with
the pressure jump pin, FP32, Poiseuille + 5% seeded noise initial state
(synthetic; bulk speed 0.05 lattice units, nu = 0.05*2R/5300 for bulk
Re ~ 5300, Blasius-estimated pin).
One step = one FSLBM time step over all 67,108,864 lattice cells.

Printed: one JSON line with MLUPS (all lattice cells, the paper's
convention), the HBM roofline of the step kernel and the CPU baseline
(the C oracle with OpenMP on a bounded sample of the same workload).

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
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain device copy of 1 GiB x 10 (the "
                         "scale for the FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} "
              f"(WORLD_SIZE={world}) is not supported. Independent pipes or "
              f"requests shard one per GPU; a single domain would need a "
              f"z-slab halo exchange that is not part of this build.",
              file=sys.stderr)
        return 2

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

    s = Solver(mask.shape, prm, mask=mask, precision=args.precision, kernel=kern)
    if args.chunk > 0:
        s.set_tile_chunk(args.chunk)
    elif args.chunk < 0:  # autotune the schedule (state is not advanced)
        s.autotune(5)
    if args.precision == "f64":
        s.set_state(*st)
    else:
        s.set_state(*[a.astype(np.float32) for a in st])
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
    achieved = B * n / (kernel_ms * 1e-3) / 1e9
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
                     "bytes_per_cell": B, "kernel_ms": round(kernel_ms, 4)},
        "mflups_fluid": round(nf * nz * args.steps / wall / 1e6, 1),
    }
    s.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mask, st, pin, nu, args.cpu_steps,
                                           args.precision)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
