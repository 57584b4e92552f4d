"""GPU tests of the file-level drop-in fslbm_run (the replacement of the
reference's main(), FSLBM.cu:489-909): it reads pipe.txt / param.txt as
pipe.m writes them, and writes qp.txt, the averaged YZ-plane statistics,
the instantaneous planes, IO.txt and the restart backups in the
reference's formats, on the reference's cadences.  The numbers it writes
are checked against the same schedule run through the library."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from flowcontrast_amd import KERNEL_AUTO, RunParams, Solver, inputs

pytestmark = pytest.mark.gpu

CLI = Path(__file__).resolve().parent.parent / "flowcontrast_amd" / "fslbm_run"
DIAM, LENGTH, FRAME = 20, 40, 4
LX, LZ = DIAM + 2 * FRAME, LENGTH + 2 * FRAME
PARAMS = (1.0, 4e-3, 1.2e-2, 2, 1, 0.0)  # tblend pin nu d1 delta blend


def write_inputs(d: Path):
    padded, _ = inputs.pipe_m_mask(DIAM, LENGTH, FRAME)
    (d / "pipe.txt").write_text("\n".join(str(int(v)) for v in padded.reshape(-1)) + "\n")
    tb, pin, nu, d1, delta, blend = PARAMS
    (d / "param.txt").write_text(f"{tb:e} {pin:e} {nu:e} {d1} {delta} {blend:e}\nrun\n")
    return padded


def run_cli(d: Path, *args):
    if not CLI.exists():
        pytest.fail(f"{CLI} is not built")
    cmd = [str(CLI), "--lx", str(LX), "--ly", str(LX), "--lz", str(LZ), "--d2", str(FRAME),
           "--quiet", *map(str, args)]
    res = subprocess.run(cmd, cwd=d, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    return res


def library_run(padded, steps, monitor, stats, precision):
    """The CLI's schedule through the Python binding of the same library."""
    folded = inputs.fold_mask(padded, FRAME)
    nz = folded.shape[0]
    tb, pin, nu, d1, delta, blend = (np.float32(v) if isinstance(v, float) else v
                                     for v in PARAMS)
    prm = RunParams(tblend=float(tb), pin=float(pin), nu=float(nu), d1=d1, delta=delta,
                    blend=float(blend))
    p = np.where(folded == 0, float(pin) * (nz - np.arange(nz))[:, None, None] / nz, 0.0)
    z = np.zeros_like(p)
    qp = []
    with Solver(folded.shape, prm, mask=folded, precision=precision, kernel=KERNEL_AUTO) as s:
        s.set_state(p, z, z, z)
        s.stats_config(LX // 2, FRAME)
        for c in range(1, steps + 1):
            s.step(1)
            if c % monitor:
                continue
            qp.append(s.monitor(True))
            if c % stats == 0:
                s.stats_accumulate(delta)
        acc, cnt, cnt2 = s.stats_read()
        planes = [s.plane(0, LX // 2, FRAME, f) for f in range(4)]
    return np.array(qp), acc, cnt, cnt2, planes


def read_col(path):
    return np.loadtxt(path, ndmin=1)


def test_cli_files_match_the_library(tmp_path):
    padded = write_inputs(tmp_path)
    run_cli(tmp_path, "--steps", 100, "--monitor", 50, "--stats", 50, "--backup", 0,
            "--precision", "f32", "--fresh")
    qp_lib, acc, cnt, cnt2, planes = library_run(padded, 100, 50, 50, "f32")

    qp = np.loadtxt(tmp_path / "qp.txt", ndmin=2)
    assert qp.shape == (2, 6)
    # q1 q2 p1 p2 as the library's monitor returns them (written with %e)
    np.testing.assert_allclose(qp[:, :4], qp_lib, rtol=2e-6, atol=1e-12)
    # the outlet plane is the image of the inlet plane: p1 - p2 = pin, q1 = q2
    np.testing.assert_allclose(qp[:, 2] - qp[:, 3], PARAMS[1], rtol=1e-3)
    np.testing.assert_allclose(qp[:, 0], qp[:, 1], rtol=1e-5)

    npts = LX * LZ
    names = ["_p", "_ux", "_uy", "_uz", "_p2", "_ux2", "_uy2", "_uz2", "_uxy2", "_uxz2", "_uyz2"]
    for k, name in enumerate(names):
        got = read_col(tmp_path / f"{name}.txt")
        assert got.shape == (npts,)
        exp = acc[k] / (cnt if k < 4 else cnt2)
        np.testing.assert_allclose(got, exp.astype(np.float32), rtol=2e-6, atol=1e-30)
    assert (tmp_path / "_av.txt").read_text().split() == [str(cnt), str(cnt2)]

    for f, name in enumerate(["p.txt", "ux.txt", "uy.txt", "uz.txt"]):
        got = read_col(tmp_path / name)
        np.testing.assert_allclose(got, planes[f].astype(np.float32), rtol=2e-6, atol=1e-30)
    io = read_col(tmp_path / "IO.txt").astype(int)
    assert np.array_equal(io, padded[:, :, LX // 2].reshape(-1))


def test_cli_backup_and_restart(tmp_path):
    write_inputs(tmp_path)
    run_cli(tmp_path, "--steps", 100, "--monitor", 50, "--stats", 50, "--backup", 100,
            "--precision", "f32", "--fresh")
    n = LX * LX * LZ
    for v in ("p3Dii", "ux3Dii", "uy3Dii", "uz3Dii"):
        assert (tmp_path / v).stat().st_size == 4 * n
    vol = np.fromfile(tmp_path / "uz3Dii", dtype=np.float32).reshape(LZ, LX, LX)
    assert np.isfinite(vol).all() and np.abs(vol).max() > 0
    cnt0 = (tmp_path / "_avii.txt").read_text().split()
    # restart: the volumes and averages are loaded, qp.txt continues
    run_cli(tmp_path, "--steps", 50, "--monitor", 50, "--stats", 50, "--backup", 0,
            "--precision", "f32")
    qp = np.loadtxt(tmp_path / "qp.txt", ndmin=2)
    assert qp.shape == (3, 6)
    cnt1 = (tmp_path / "_av.txt").read_text().split()
    assert int(cnt1[0]) == int(cnt0[0]) + 1


@pytest.mark.parametrize("precision", ["f64", "f32", "f16s", "bf16s"])
def test_cli_exact_checkpoint_resumes_bitwise(tmp_path, precision):
    """A run interrupted at a backup and resumed from fslbm_ckpt.bin writes
    the same files as an uninterrupted run (state, statistics sums, step
    number and pending gauge are restored exactly; in the 16-bit storage
    modes also the stored values, their scales and the rounding counter)."""
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    write_inputs(a)
    write_inputs(b)
    common = ["--monitor", 50, "--stats", 50, "--precision", precision]
    run_cli(a, "--steps", 150, "--backup", 0, "--fresh", *common)
    run_cli(b, "--steps", 100, "--backup", 100, "--fresh", *common)
    assert (b / "fslbm_ckpt.bin").exists()
    run_cli(b, "--steps", 50, "--backup", 0, *common)
    for name in ("qp.txt", "_p.txt", "_uz.txt", "_uz2.txt", "_uxz2.txt", "_av.txt", "p.txt",
                 "uz.txt", "uz1.txt"):
        assert (a / name).read_text() == (b / name).read_text(), name
