"""GPU tests of the file-level drop-in fslbm_run (the replacement of the
reference's main(), FSLBM.cu:489-909): it reads pipe.txt / param.txt as
pipe.m writes them, and writes qp.txt, the averaged YZ-plane statistics,
the instantaneous planes, IO.txt and the restart backups in the
reference's formats, on the reference's cadences.  The numbers it writes
are checked against the same schedule run through the library, and against
the CPU oracle's run of the reference schedule in the reference's padded
layout (monitoring, statistics, planes, restart volumes and both restart
paths)."""
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


# ----------------------------------------------------------------------
# the files against the CPU oracle in the reference's padded layout
# ----------------------------------------------------------------------
def param_values():
    """param.txt as the CLI reads it: %e into float (FSLBM.cu:502-507)."""
    tb, pin, nu, d1, delta, blend = PARAMS
    return (float(np.float32(tb)), float(np.float32(pin)), float(np.float32(nu)), d1, delta,
            float(np.float32(blend)))


def oracle_schedule(orc, padded, steps, monitor, stats, dtype, state=None, acc=None, c0=0,
                    p2=0.0):
    """The reference's main loop (FSLBM.cu:620-764) on the padded layout:
    FSLBMcalc + bcond with the pending p2, the inlet/outlet monitoring every
    `monitor` steps (setting p2), and every `stats` steps the YZ-plane
    statistics at x = lx/2.  Returns the qp rows, the statistics
    accumulator, the state, the planes of the last statistics step and the
    pending p2.  c0 and p2 continue an earlier run (its step count and the
    gauge its last monitoring step left for the next step)."""
    tb, pin, nu, d1, delta, blend = param_values()
    op = orc.params(tb, pin, nu, blend, d1)
    lz, ly, lx = padded.shape
    if state is None:  # initphys: p = pin (lz - d2 - z) / (lz - 2 d2), u = 0
        z = np.arange(lz)[:, None, None]
        p = np.where(padded != 125, pin * (lz - FRAME - z) / (lz - 2 * FRAME), 0.0)
        state = [p.astype(dtype)] + [np.zeros(padded.shape, dtype) for _ in range(3)]
    if acc is None:
        acc = orc.Stats(ly * lz)
    l1 = float((padded[FRAME] == 0).sum())
    r = (lx - 2 * FRAME) // 2
    qp, planes = [], None
    for c in range(c0 + 1, c0 + steps + 1):
        state = orc.step_padded(state, op, padded, FRAME, p2=p2, dtype=dtype)
        p2 = 0.0
        if c % monitor:
            continue
        q1, q2, p1, p2 = orc.monitor_padded(state, FRAME, l1)
        ff = (p1 - p2) / q1 / q1 / 3.0 * 4.0 * r / (lz - 2 * FRAME)
        qp.append([q1, q2, p1, p2, ff, float(state[3][lz // 2, ly // 2, lx // 2])])
        if c % stats == 0:
            yz = [np.asarray(f, np.float64)[:, :, lx // 2].reshape(-1) for f in state]
            acc.accumulate(yz, delta)
            planes = yz + [np.asarray(f, np.float64)[lz // 2].reshape(-1) for f in state]
    return np.array(qp), acc, state, planes, p2


def check_files(d: Path, qp_ref, acc, planes, rtol, atol_rel, rows=None):
    qp = np.loadtxt(d / "qp.txt", ndmin=2)
    if rows is not None:
        qp = qp[rows]
    assert qp.shape == qp_ref.shape
    for k in range(6):
        # q1, q2 share the velocity scale and p1, p2 the pressure scale (p2
        # is zero up to roundoff once the gauge has been applied)
        cols = {0: [0, 1], 1: [0, 1], 2: [2, 3], 3: [2, 3]}.get(k, [k])
        scale = np.abs(qp_ref[:, cols]).max()
        np.testing.assert_allclose(qp[:, k], qp_ref[:, k], rtol=rtol, atol=atol_rel * scale,
                                   err_msg=f"qp.txt column {k}")
    a, cnt, cnt2 = acc.read()
    assert (d / "_av.txt").read_text().split() == [str(cnt), str(cnt2)]
    names = ["_p", "_ux", "_uy", "_uz", "_p2", "_ux2", "_uy2", "_uz2", "_uxy2", "_uxz2", "_uyz2"]
    exp = [a[k] / (cnt if k < 4 else cnt2) for k in range(11)]
    # Scales: the velocity components share one (in this axisymmetric flow
    # ux and uy are roundoff-sized); a deviation takes its mean field's
    # scale (the pressure is stationary here, so |dev p| is roundoff), and
    # a product of deviations the velocity scale squared.
    ps = np.abs(exp[0]).max()
    us = max(np.abs(exp[j]).max() for j in (1, 2, 3, 5, 6, 7))
    scales = [ps, us, us, us, ps, us, us, us, us * us, us * us, us * us]
    for k, name in enumerate(names):
        scale = scales[k]
        np.testing.assert_allclose(read_col(d / f"{name}.txt"), exp[k], rtol=rtol,
                                   atol=atol_rel * scale + 1e-30, err_msg=name)
    files = ["p.txt", "ux.txt", "uy.txt", "uz.txt", "p1.txt", "ux1.txt", "uy1.txt", "uz1.txt"]
    for k, (name, e) in enumerate(zip(files, planes)):
        js = [k] if k % 4 == 0 else [4 * (k // 4) + j for j in (1, 2, 3)]
        scale = max(np.abs(planes[j]).max() for j in js)
        np.testing.assert_allclose(read_col(d / name), e, rtol=rtol,
                                   atol=atol_rel * scale + 1e-30, err_msg=name)


# FP64: the files hold %e text (7 significant digits) of values that agree
# with the oracle to ~1e-12; FP32: relative 1e-5 of each field's scale.
ORACLE_TOL = {"f64": (2e-6, 1e-9), "f32": (1e-5, 1e-5)}


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_cli_files_match_padded_oracle(tmp_path, orc, precision):
    """fslbm_run's qp.txt, averaged statistics, instantaneous YZ/XY planes
    and restart volumes against the oracle's padded-layout run of the
    reference schedule; then both restart paths continue it: the exact
    checkpoint and the reference's float32 volumes with %e averages."""
    rtol, atol = ORACLE_TOL[precision]
    dtype = np.float64 if precision == "f64" else np.float32
    padded = write_inputs(tmp_path)
    run_cli(tmp_path, "--steps", 100, "--monitor", 50, "--stats", 50, "--backup", 100,
            "--precision", precision, "--fresh")
    qp_ref, acc, state, planes, p2 = oracle_schedule(orc, padded, 100, 50, 50, dtype)
    check_files(tmp_path, qp_ref, acc, planes, rtol, atol)
    lz, ly, lx = padded.shape
    uscale = max(np.abs(np.asarray(state[f], np.float64)).max() for f in (1, 2, 3))
    for f, v in enumerate(("p3Dii", "ux3Dii", "uy3Dii", "uz3Dii")):
        vol = np.fromfile(tmp_path / v, dtype=np.float32).reshape(lz, ly, lx)
        ref = np.asarray(state[f], np.float64)
        scale = np.abs(ref).max() if f == 0 else uscale
        np.testing.assert_allclose(vol, ref.astype(np.float32), rtol=max(rtol, 1e-6),
                                   atol=max(atol, 1e-7) * scale + 1e-30, err_msg=v)

    # (1) exact checkpoint: 50 more steps continue the oracle's run
    exact = tmp_path / "exact"
    exact.mkdir()
    for name in tmp_path.iterdir():
        if name.is_file():
            (exact / name.name).write_bytes(name.read_bytes())
    run_cli(exact, "--steps", 50, "--monitor", 50, "--stats", 50, "--backup", 0,
            "--precision", precision)
    acc1 = orc.Stats(ly * lz)
    a0, c0, c20 = acc.read()
    acc1.acc[:] = a0.reshape(-1)
    acc1.cnt.value, acc1.cnt2.value = c0, c20
    qp1, acc1, _, planes1, _ = oracle_schedule(orc, padded, 50, 50, 50, dtype, state=state,
                                               acc=acc1, c0=100, p2=p2)
    check_files(exact, np.vstack([qp_ref, qp1]), acc1, planes1, rtol, atol)

    # (2) the reference's restart files: float32 volumes and %e averages
    # (FSLBM.cu:386-468); the oracle restarts from the same float32 values
    vols = tmp_path / "vols"
    vols.mkdir()
    for name in tmp_path.iterdir():
        if name.is_file() and name.name != "fslbm_ckpt.bin":
            (vols / name.name).write_bytes(name.read_bytes())
    run_cli(vols, "--steps", 50, "--monitor", 50, "--stats", 50, "--backup", 0,
            "--precision", precision)
    st32 = [np.fromfile(vols / v, dtype=np.float32).reshape(lz, ly, lx).astype(dtype)
            for v in ("p3Dii", "ux3Dii", "uy3Dii", "uz3Dii")]
    acc2 = orc.Stats(ly * lz)
    cnt, cnt2 = map(int, (vols / "_avii.txt").read_text().split())
    names = ["_p", "_ux", "_uy", "_uz", "_p2", "_ux2", "_uy2", "_uz2", "_uxy2", "_uxz2", "_uyz2"]
    acc2.acc[:] = np.concatenate([
        read_col(vols / f"{n}ii.txt").astype(np.float32).astype(np.float64)
        * (cnt if k < 4 else cnt2) for k, n in enumerate(names)])
    acc2.cnt.value, acc2.cnt2.value = cnt, cnt2
    # the reference restarts with no pending gauge (p2 is not saved)
    qp2, acc2, _, planes2, _ = oracle_schedule(orc, padded, 50, 50, 50, dtype, state=st32,
                                               acc=acc2, c0=100)
    check_files(vols, qp2, acc2, planes2, max(rtol, 1e-5), max(atol, 1e-6), rows=[2])
