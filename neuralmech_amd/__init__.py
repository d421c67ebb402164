"""MI355X-native finite-difference scalar-wave solver (gfx950 HIP kernels behind a C ABI).

The hot path is the leapfrog step of ``projects/18_simple_wave_fd`` and the wave solver that the
``projects/18_fwi`` and transient topology-optimisation drivers call; see ``include/wavefd.h``.
"""

from . import _lib
from .solver import GradientPlan, Medium, PointSet, WaveGrid, run, stable_dt

__all__ = ["GradientPlan", "Medium", "PointSet", "WaveGrid", "run", "stable_dt", "_lib"]
