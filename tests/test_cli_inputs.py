"""CPU checks of the drop-in driver fslbm_run's input handling.

The reference reads param.txt and pipe.txt without checking fopen or
fscanf results, and its read_geom reads one value past the mask
(FSLBM.cu:478-485,499-510).  The driver validates every input before it
touches a device: a missing or truncated file ends the run with a message
and exit status 1, so these run without a GPU."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from flowcontrast_amd import inputs

CLI = Path(__file__).resolve().parent.parent / "flowcontrast_amd" / "fslbm_run"
DIAM, LENGTH, FRAME = 8, 12, 2
LX, LZ = DIAM + 2 * FRAME, LENGTH + 2 * FRAME


def run(d: Path):
    if not CLI.exists():
        pytest.skip("fslbm_run is not built")
    cmd = [str(CLI), "--lx", str(LX), "--ly", str(LX), "--lz", str(LZ), "--d2", str(FRAME),
           "--steps", "1", "--quiet", "--fresh"]
    return subprocess.run(cmd, cwd=d, capture_output=True, text=True, timeout=60)


def write(d: Path, pipe=True, param=True, cut_pipe=0, param_text=None, code=None):
    padded, _ = inputs.pipe_m_mask(DIAM, LENGTH, FRAME)
    vals = [str(int(v)) for v in padded.reshape(-1)]
    if code is not None:
        vals[len(vals) // 2] = str(code)
    if cut_pipe:
        vals = vals[:-cut_pipe]
    if pipe:
        (d / "pipe.txt").write_text("\n".join(vals) + "\n")
    if param:
        (d / "param.txt").write_text(param_text or "1.0e+00 3.0e-03 4.0e-03 2 1 0.0e+00\n")


@pytest.mark.parametrize("kwargs, needle", [
    (dict(param=False), "param.txt"),
    (dict(pipe=False), "pipe.txt"),
    (dict(cut_pipe=5), "pipe.txt"),
    (dict(param_text="1.0 3e-3 4e-3\n"), "param.txt"),
    (dict(code=7), "mask codes"),
    (dict(code=381), "mask codes"),
])
def test_bad_inputs_fail_with_a_message(tmp_path, kwargs, needle):
    write(tmp_path, **kwargs)
    res = run(tmp_path)
    assert res.returncode == 1, (res.returncode, res.stderr)
    assert needle in res.stderr, res.stderr
    assert not (tmp_path / "qp.txt").exists()


def test_padding_must_be_periodic(tmp_path):
    padded, _ = inputs.pipe_m_mask(DIAM, LENGTH, FRAME)
    padded = padded.copy()
    padded[0, LX // 2, LX // 2] = 125  # a wall in one padding plane only
    (tmp_path / "pipe.txt").write_text("\n".join(str(int(v)) for v in padded.reshape(-1)))
    (tmp_path / "param.txt").write_text("1.0e+00 3.0e-03 4.0e-03 2 1 0.0e+00\n")
    res = run(tmp_path)
    assert res.returncode == 1 and "periodic" in res.stderr, res.stderr
    assert np.any(padded == 1)
