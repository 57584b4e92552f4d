"""flowcontrast_amd: MI355X (gfx950) FSLBM time step.

A drop-in for the hot path of alparmat/FlowContrast
(PipeDemoStandalone/FSLBM.cu): the fractional-step lattice Boltzmann step
(D3Q27 equilibrium predictor, CCS reverse-diffusion corrector, blend),
its wall and axial-periodic pressure-jump boundaries, and the mean-field /
fluctuation statistics, computed by hand-written HIP kernels behind the C
ABI in include/fslbm.h.
"""
from ._lib import (BF16S, F16S, F32, F64, KERNEL_AUTO, KERNEL_PLAIN,  # noqa: F401
                   KERNEL_TILED, FslbmError, lib)
from .solver import RunParams, Solver, tile_plan  # noqa: F401
from . import inputs  # noqa: F401
from .batch import Batch, BudgetFull  # noqa: F401
from .serve import Completion, Server  # noqa: F401
