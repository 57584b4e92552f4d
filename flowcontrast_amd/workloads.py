"""Seeded solve workloads for batches and the serving loop.

* uneven_pipes: independent periodic pipes of uneven radius, length and
  bulk Reynolds number (radius U[12, 48] cells, length U[4, 10] diameters,
  Re U[200, 3000]), Poiseuille + 2% noise initial states, FP32.
* serving_requests: a Poisson stream of solve requests, 80% pipes (radius
  U[16, 64], length U[4, 8] diameters, Re U[200, 6000]) and 20%
  Taylor-Green boxes (side U[64, 192], Re U[100, 1600]), each with a step
  budget U[2000, 20000] and FP32 or FP16-shifted storage.

Pipes run at bulk speed 0.05 (nu = 0.05 * 2R / Re) with the laminar jump
pin for that speed; Taylor-Green boxes at U0 = 0.05.  Each case builds
its mask and state on demand, from its own seed, so a workload of many
large domains never has to sit in host memory at once.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import inputs
from .solver import RunParams

UBAR = 0.05


@dataclass
class Case:
    kind: str          # "pipe" or "tgv"
    radius: float      # pipes: radius; boxes: 0
    length: int        # z cells
    side: int          # x/y cells
    re: float
    steps: int
    precision: str
    seed: int
    noise: float = 0.02
    arrival: float = 0.0  # seconds after the stream starts

    @property
    def shape(self):
        return (self.length, self.side, self.side)

    @property
    def cells(self) -> int:
        return self.length * self.side * self.side

    def build(self):
        """(RunParams, mask or None, [p, ux, uy, uz]) of the case."""
        if self.kind == "tgv":
            n = self.side
            nu = inputs.tgv_viscosity(n, UBAR, self.re)
            prm = RunParams(tblend=1.0, pin=0.0, nu=nu, d1=2, delta=1, blend=0.0)
            return prm, None, inputs.taylor_green_state(n, UBAR, 1.0)
        R = self.radius
        mask, _ = inputs.circle_mask(self.side, self.side, self.length, R)
        nu = UBAR * 2 * R / self.re
        pin = inputs.laminar_pin(nu, self.length, UBAR, R)
        st = inputs.poiseuille_state(mask, R, UBAR, pin, self.noise, self.seed)
        prm = RunParams(tblend=1.0, pin=pin, nu=nu, d1=2, delta=1, blend=0.0)
        return prm, mask, st


def pipe_side(radius: float) -> int:
    """Cross-section side for a centred pipe with a wall frame of at
    least one cell on every side."""
    return int(math.ceil(2 * radius)) + 4


def uneven_pipes(count: int = 64, seed: int = 4, steps: int = 10000):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        R = float(rng.uniform(12.0, 48.0))
        k = float(rng.uniform(4.0, 10.0))
        re = float(rng.uniform(200.0, 3000.0))
        out.append(Case("pipe", R, int(round(k * 2 * R)), pipe_side(R), re, steps, "f32",
                        seed * 100000 + i))
    return out


def serving_requests(count: int = 32, seed: int = 5, rate: float = 2.0,
                     fp16_share: float = 0.5):
    """`count` requests with Poisson arrivals at `rate` per second."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += float(rng.exponential(1.0 / rate))
        steps = int(rng.integers(2000, 20001))
        prec = "f16s" if rng.uniform() < fp16_share else "f32"
        if rng.uniform() < 0.8:
            R = float(rng.uniform(16.0, 64.0))
            k = float(rng.uniform(4.0, 8.0))
            re = float(rng.uniform(200.0, 6000.0))
            c = Case("pipe", R, int(round(k * 2 * R)), pipe_side(R), re, steps, prec,
                     seed * 100000 + i)
        else:
            n = int(rng.integers(64, 193))
            re = float(rng.uniform(100.0, 1600.0))
            c = Case("tgv", 0.0, n, n, re, steps, prec, seed * 100000 + i, noise=0.0)
        c.arrival = t
        out.append(c)
    return out
