"""Geometry and seeded initial states (host-side builders of libfslbm.so).

These wrap the C ABI builders so the CLI, the tests and the benchmark use
one implementation: pipe.m's mask (pipe.m:38-58), a centred circular pipe
for the folded layout, a rough-walled pipe, the Poiseuille + seeded-noise
state and the Taylor-Green vortex.  Arrays are shaped (nz, ny, nx), x fastest.
"""
from __future__ import annotations

import math

import numpy as np

from ._lib import check, lib


def pipe_m_mask(diameter: int = 100, length: int = 300, frame: int = 10):
    """pipe.m's padded mask: codes 0 fluid, 1 padding, 125 wall.

    Returns (mask of shape (length+2*frame, d+2*frame, d+2*frame), number
    of padding cells)."""
    lx = diameter + 2 * frame
    lz = length + 2 * frame
    m = np.empty(lx * lx * lz, dtype=np.uint8)
    npad = check(lib().fslbm_pipe_m_mask(diameter, length, frame, m))
    return m.reshape(lz, lx, lx), npad


def fold_mask(padded, d2: int):
    """Fold a reference mask with d2 padding planes per end (0/125 codes)."""
    lz, ly, lx = padded.shape
    out = np.empty(lx * ly * (lz - 2 * d2), dtype=np.uint8)
    check(lib().fslbm_fold_mask(lx, ly, lz, d2,
                                np.ascontiguousarray(padded, dtype=np.uint8).reshape(-1),
                                out))
    return out.reshape(lz - 2 * d2, ly, lx)


def circle_mask(nx: int, ny: int, nz: int, radius: float):
    """Centred circular pipe; returns (mask, fluid cells per cross-section)."""
    m = np.empty(nx * ny * nz, dtype=np.uint8)
    nf = check(lib().fslbm_circle_mask(nx, ny, nz, float(radius), m))
    return m.reshape(nz, ny, nx), nf


def rough_pipe_mask(nx: int, ny: int, nz: int, radius: float, amp: float,
                    nbumps: int, seed: int):
    """A rough-walled pipe, a mask that varies along z (the method paper
    validates rough pipes, MethodDoc_withFSLBMpaper.pdf p.6): a centred pipe
    whose wall radius R + amp sin(3 theta + 2 pi z/nz) cos(4 pi z/nz) varies
    with angle and z, plus `nbumps` seeded spherical bumps of radius amp to
    2 amp centred on the nominal wall (periodic in z), and a 2-cell wall
    frame on the x/y faces.  Returns (mask, number of fluid cells)."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:ny, 0:nx].astype(np.float64)
    cx, cy = (nx - 1) / 2.0, (ny - 1) / 2.0
    r = np.hypot(x - cx, y - cy)
    th = np.arctan2(y - cy, x - cx)
    frame = np.zeros((ny, nx), dtype=bool)
    frame[:, :2] = frame[:, -2:] = True
    frame[:2, :] = frame[-2:, :] = True
    mask = np.empty((nz, ny, nx), dtype=np.uint8)
    for z in range(nz):
        rw = radius + amp * np.sin(3 * th + 2 * math.pi * z / nz) * math.cos(4 * math.pi * z / nz)
        mask[z] = np.where((r > rw) | frame, 125, 0)
    for _ in range(nbumps):
        t = rng.uniform(0, 2 * math.pi)
        zc = rng.uniform(0, nz)
        rb = rng.uniform(amp, 2 * amp)
        bx, by = cx + radius * math.cos(t), cy + radius * math.sin(t)
        xs = slice(max(0, int(bx - rb)), min(nx, int(bx + rb) + 2))
        ys = slice(max(0, int(by - rb)), min(ny, int(by + rb) + 2))
        d2 = (x[ys, xs] - bx) ** 2 + (y[ys, xs] - by) ** 2
        for z in range(math.floor(zc - rb), math.ceil(zc + rb) + 1):
            h2 = rb * rb - (z - zc) ** 2
            if h2 >= 0:
                mask[z % nz, ys, xs][d2 <= h2] = 125
    return mask, int(np.count_nonzero(mask == 0))


def poiseuille_state(mask, radius: float, ubar: float, pin: float,
                     noise: float, seed: int):
    """initphys pressure ramp + Poiseuille profile + seeded noise."""
    nz, ny, nx = mask.shape
    n = nx * ny * nz
    arrs = [np.empty(n) for _ in range(4)]
    check(lib().fslbm_init_poiseuille(nx, ny, nz, np.ascontiguousarray(mask).reshape(-1),
                                      float(radius), float(ubar), float(pin),
                                      float(noise), int(seed), *arrs))
    return [a.reshape(nz, ny, nx) for a in arrs]


def taylor_green_state(n: int, u0: float = 0.05, rho0: float = 1.0):
    """Taylor-Green vortex on an n^3 periodic box (stored sign)."""
    arrs = [np.empty(n ** 3) for _ in range(4)]
    check(lib().fslbm_init_taylor_green(n, n, n, float(u0), float(rho0), *arrs))
    return [a.reshape(n, n, n) for a in arrs]


def laminar_pin(nu: float, length: int, ubar: float, radius: float) -> float:
    """Jump pin for a laminar mean speed ubar: ubar = pin R^2 / (24 nu L)
    with P = p/3 (Hagen-Poiseuille in lattice units)."""
    return 24.0 * nu * length * ubar / (radius * radius)


def blasius_pin(nu: float, length: int, re_bulk: float, radius: float) -> float:
    """Jump pin from the Blasius friction factor f = 0.316 Re^-1/4 with the
    reference's f = 4 r dp / (3 ubar^2 L) (FSLBM.cu:660)."""
    ubar = re_bulk * nu / (2.0 * radius)
    f = 0.316 * re_bulk ** -0.25
    return 3.0 * f * ubar * ubar * length / (4.0 * radius)


def tgv_viscosity(n: int, u0: float = 0.05, re: float = 1600.0) -> float:
    """Lattice viscosity of a Taylor-Green run: Re = U0 (n / 2 pi) / nu."""
    return u0 * (n / (2.0 * math.pi)) / re
