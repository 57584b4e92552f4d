"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer:
every entry point on small masked, periodic and padded cases (both
precisions, d1 1 to 3, blend 0 and 1, gauge, wrapped z), in a child process
with the sanitizer runtime preloaded and every report fatal."""
import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

SCRIPT = r'''
import numpy as np
from oracle import oracle as O
from flowcontrast_amd import inputs
for (nx, ny, nz, R) in [(12, 11, 9, 4.0), (9, 9, 5, 3.0)]:
    mask, _ = inputs.circle_mask(nx, ny, nz, R)
    st = inputs.poiseuille_state(mask, R, 0.03, 2e-3, 0.05, 1)
    for d1 in (1, 2, 3):
        for b in (0.0, 1.0):
            for dt in (np.float64, np.float32):
                p = O.params(0.9, 2e-3, 0.02, b, d1)
                O.step_folded(st, p, mask=mask, gauge=1e-4, nsteps=2, dtype=dt)
                O.step_folded(st, p, nsteps=2, dtype=dt)
pad, _ = inputs.pipe_m_mask(8, 10, 3)
lz, ly, lx = pad.shape
z = np.arange(lz)[:, None, None]
s = [np.where(pad != 125, 1e-3 * (lz - 3 - z) / (lz - 6), 0.0)] + [np.zeros(pad.shape)] * 3
for dt in (np.float64, np.float32):
    s2 = O.step_padded(s, O.params(1.0, 1e-3, 0.02, 0.0, 2), pad, 3, p2=1e-4, dtype=dt)
O.monitor_padded(s2, 3, float((pad[3] == 0).sum()))
acc = O.Stats(ly * lz)
for delta in (2, 1, 0, 2):
    acc.accumulate([a[:, :, lx // 2].reshape(-1) for a in s2], delta)
print("sanitized oracle ok")
'''


def test_oracle_under_sanitizers(tmp_path):
    gcc = shutil.which("gcc")
    if not gcc:
        pytest.skip("gcc not available")
    res = subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "sanitize"],
                         capture_output=True, text=True)
    if res.returncode != 0:
        pytest.skip("sanitizer build unavailable: " + res.stderr[-300:])
    libs = [subprocess.run([gcc, f"-print-file-name={n}"], capture_output=True,
                           text=True).stdout.strip() for n in ("libasan.so", "libubsan.so")]
    if not all(os.path.isabs(x) and Path(x).exists() for x in libs):
        pytest.skip("sanitizer runtimes not found")
    # the sanitizer runtimes go first; whatever the environment already
    # preloads stays preloaded after them
    pre = [x for x in os.environ.get("LD_PRELOAD", "").split(":") if x]
    env = dict(os.environ, LD_PRELOAD=":".join(libs + pre),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               FSLBM_ORACLE_SO=str(ROOT / "oracle" / "_build" / "liboracle_san.so"),
               OMP_NUM_THREADS="2", PYTHONPATH=str(ROOT))
    run = subprocess.run([sys.executable, "-c", SCRIPT], capture_output=True, text=True,
                         env=env, cwd=str(ROOT), timeout=600)
    assert run.returncode == 0, run.stderr[-3000:]
    assert "sanitized oracle ok" in run.stdout
    assert "runtime error" not in run.stderr and "AddressSanitizer" not in run.stderr
