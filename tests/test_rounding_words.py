"""Statistics of the stochastic-rounding words of the 16-bit storage modes.

The kernels round each stored 16-bit value stochastically with a random
operand per field, cell and step (fslbm_store.h: sr_salt, sr_plane,
sr_words).  The conversion takes its random bits from the top of the
operand (13 bits for binary16, 16 for bfloat16).  Rounding is unbiased when, for
every cell, these 16-bit draws are uniform and independent from step to
step; neighbouring cells and the four fields of one cell should draw
unrelated values, so that rounding errors do not line up in space.

A host harness built from the kernels' own header writes the words; numpy
measures the mean, a 16-bin histogram, the correlation between
neighbouring cells along x, y and z, between the four fields and between
consecutive steps, and the rate at which a residual r rounds up.  A
generator of the form (cell ^ key) * M alone (no per-column salt) fails the
neighbour test with a correlation near -0.5; the salt removes it.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "flowcontrast_amd" / "csrc"
HARNESS = ROOT / "_build" / "rounding_words"
NX, NY, NZ, STEPS = 64, 48, 40, 4


@pytest.fixture(scope="module")
def draws():
    src = Path(__file__).resolve().parent / "rounding_words_main.cpp"
    HARNESS.parent.mkdir(exist_ok=True)
    if not HARNESS.exists() or HARNESS.stat().st_mtime < max(
            src.stat().st_mtime, (CSRC / "fslbm_store.h").stat().st_mtime):
        subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__",
                        f"-I{CSRC}", str(src), "-o", str(HARNESS)], check=True,
                       capture_output=True)
    raw = subprocess.run([str(HARNESS), str(NX), str(NY), str(NZ), str(STEPS), "11"],
                         check=True, capture_output=True).stdout
    w = np.frombuffer(raw, dtype="<u4").reshape(STEPS, NZ, NY, NX, 4)
    # the top 16 bits each field's conversion sees, as uniforms in [0, 1)
    return np.stack([(w[..., f] >> 16).astype(np.float64) / 65536.0 for f in range(4)])


def corr(u, v):
    u = u.ravel() - u.mean()
    v = v.ravel() - v.mean()
    return float((u * v).sum() / np.sqrt((u * u).sum() * (v * v).sum()))


def test_uniform(draws):
    n = draws[0].size
    for u in draws:
        assert abs(u.mean() - 0.5) < 5 * 0.2887 / np.sqrt(n)
        hist, _ = np.histogram(u, bins=16, range=(0, 1))
        chi2 = float(((hist - n / 16) ** 2 / (n / 16)).sum())
        assert chi2 < 60, chi2  # 15 degrees of freedom: p < 1e-6 above 60


def test_neighbours_fields_and_steps_are_unrelated(draws):
    lim = 5 / np.sqrt(draws[0].size)
    for u in draws:
        assert abs(corr(u[:, :, :, 1:], u[:, :, :, :-1])) < lim  # x
        assert abs(corr(u[:, :, 1:], u[:, :, :-1])) < lim        # y
        assert abs(corr(u[:, 1:], u[:, :-1])) < lim              # z
        assert abs(corr(u[1:], u[:-1])) < lim                    # next step
        # products too: a rigid offset U2 = U1 + c (mod 1) shows here
        d = (u[:, :, :, 1:] - u[:, :, :, :-1]) % 1.0
        hist, _ = np.histogram(d, bins=16, range=(0, 1))
        assert hist.max() < 1.2 * d.size / 16
    for i in range(4):
        for j in range(i + 1, 4):
            assert abs(corr(draws[i], draws[j])) < lim


@pytest.mark.parametrize("r", [0.03, 0.37, 0.5, 0.81])
def test_rounding_up_rate_equals_the_residual(draws, r):
    # stochastic rounding of a value r ulp above a representable one
    # rounds up with probability r: U < r
    for u in draws:
        rate = float((u < r).mean())
        assert abs(rate - r) < 5 * np.sqrt(r * (1 - r) / u.size)
