"""CPU tests of bench.py's contract: one GPU only (a clear message and
exit code 2 for more), the algorithmic byte counts of the roofline, the
measured-traffic lookup and the spot-check error metric."""
import json
import os
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402


def run_bench(args, env_extra=None):
    env = dict(os.environ)
    env.update(env_extra or {})
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True,
                          text=True, timeout=120, env=env)


def test_more_than_one_gpu_is_refused():
    r = run_bench(["--gpus", "2"])
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 2" in r.stderr
    assert r.stdout == ""
    r = run_bench([], {"WORLD_SIZE": "4"})
    assert r.returncode == 2 and "WORLD_SIZE=4" in r.stderr


def test_algorithmic_bytes():
    # state (4 fields) read once and written once, plus the 1-byte mask
    assert bench.algorithmic_bytes_per_cell("f32", True) == 33
    assert bench.algorithmic_bytes_per_cell("f32", False) == 32
    assert bench.algorithmic_bytes_per_cell("f16s", True) == 17
    assert bench.algorithmic_bytes_per_cell("bf16s", False) == 16
    assert bench.algorithmic_bytes_per_cell("f64", True) == 65


def test_traffic_lookup_matches_profiles():
    d = json.loads((ROOT / "profiles" / "traffic.json").read_text())
    for key, rec in d.items():
        workload, prec = key.split(":")
        got = bench.load_traffic(workload, prec)
        assert got == rec["bytes_per_launch"]
        # The algorithmic bytes count the cells of the stepped tiles; in a
        # walled box with a z-invariant mask the kernel neither reads nor
        # writes those tiles' quarter-rows of wall cells (17.9% of the cells
        # on the R 120 pipe), so the measured bytes may fall below them, but
        # not by more than that share.
        assert got >= 0.8 * rec["algorithmic_bytes"]
        # the writes: the new state of the same cells, less those quarter-rows
        state = {"f64": 32, "f32": 16, "f16s": 8, "bf16s": 8}[prec]
        cells = rec["algorithmic_bytes"] / bench.algorithmic_bytes_per_cell(prec, True)
        assert 0.8 * state * cells <= rec["write_bytes"] <= 1.01 * state * cells
    assert bench.load_traffic("no_such_workload", "f32") is None


def test_field_errors_metric():
    rng = np.random.default_rng(0)
    st = [1.0 + 1e-3 * rng.standard_normal((4, 5, 6))] + \
        [0.05 * rng.standard_normal((4, 5, 6)) for _ in range(3)]
    st[3] += 0.05  # a mean axial velocity
    e = bench.field_errors(st, st)
    assert e["u_rel_l2"] == 0 and e["p_rel_l2"] == 0 and e["ke_rel"] == 0
    assert e["mean_uz_rel"] == 0
    # a constant pressure offset is a gauge, not an error
    off = [st[0] + 0.3] + st[1:]
    assert bench.field_errors(off, st)["p_rel_l2"] < 1e-12
    # 1% more velocity: 1% relative L2, ~2% kinetic energy
    big = [st[0]] + [1.01 * a for a in st[1:]]
    e = bench.field_errors(big, st)
    assert abs(e["u_rel_l2"] - 0.01) < 1e-9 and abs(e["ke_rel"] - 0.0201) < 1e-9
