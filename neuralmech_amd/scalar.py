"""``ScalarWave``: the simulation descriptor of the cuwave call sites, over the wavefd solver.

Replaces ``cuwave.scalar.ScalarWave(Nx, dx, N, dt, threads, precision, space_order, wavespeed,
density)`` as used at ``projects/18_fwi/helper.py:63-72``, ``plate.py:48-49`` and
``9_structural_optimization/fwi_penalization.py:77-87``.  It is a frozen dataclass, so
``dataclasses.replace(sim, damping=sponge(...))`` works as at
``8_physics_drivers/topopt_acoustic2D.py:132-134``.

Conventions of this build (cuwave's own are not available to check against):
* ``Nx`` counts nodes per axis including the boundary ring (homogeneous Dirichlet); node ``i`` sits
  at coordinate ``(i - 1) * dx``, so the interior nodes ``1 .. n-2`` span ``[0, (n-3) dx]``, the
  ``dx = L / (n - 3)`` convention of ``helper.py:51`` and ``plate.py:40``.
* ``Nx_padded`` equals ``Nx``: node arrays have the node shape; the device layout's padding is
  internal to the solver.
* ``wavespeed`` is a number or a node array; ``density`` is accepted and does not enter the
  scalar model (``u_tt = v^2 lap u``).
* ``threads`` is accepted for compatibility; launch shapes are chosen by the solver's tuner.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Optional, Sequence

import numpy as np
import torch

from .solver import stable_dt as _stable_dt

_PRECISIONS = {"float32": torch.float32, "float64": torch.float64}


@dataclass(frozen=True)
class ScalarWave:
    Nx: Sequence[int]
    dx: Sequence[float]
    N: int
    dt: float
    threads: Optional[Sequence[int]] = None
    precision: str = "float32"
    space_order: int = 2
    wavespeed: Any = 1.0
    density: Any = 1.0
    damping: Optional[tuple] = None
    device: str = "cuda"
    _cache: dict = field(default_factory=dict, compare=False, repr=False)

    def __post_init__(self):
        object.__setattr__(self, "Nx", tuple(int(n) for n in self.Nx))
        dx = self.dx if np.ndim(self.dx) else [self.dx] * len(self.Nx)
        object.__setattr__(self, "dx", tuple(float(h) for h in dx))
        if len(self.dx) != len(self.Nx):
            raise ValueError("dx needs one spacing per axis")
        if self.precision not in _PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}")
        if self.space_order not in (2, 4, 6, 8):
            raise ValueError("space_order must be 2, 4, 6 or 8")
        vmax = float(torch.as_tensor(self.wavespeed, dtype=torch.float64).max())
        limit = _stable_dt(self.dx, vmax, self.space_order)
        if self.dt > limit * (1 + 1e-12):
            raise ValueError(f"dt = {self.dt:.6g} exceeds the stable step {limit:.6g} for wavespeed up to {vmax:.6g}")

    @property
    def ndim(self) -> int:
        return len(self.Nx)

    @property
    def dtype(self) -> torch.dtype:
        return _PRECISIONS[self.precision]

    @property
    def Nx_padded(self) -> tuple:
        return self.Nx

    @property
    def boundary(self) -> str:
        return "dirichlet"

    @property
    def cell_volume(self) -> float:
        return float(np.prod(self.dx))

    def velocity(self) -> torch.Tensor:
        """Wave speed as a float64 node array on the device."""
        v = torch.as_tensor(self.wavespeed, dtype=torch.float64)
        if v.dim() == 0:
            v = v.expand(self.Nx)
        if tuple(v.shape) != self.Nx:
            raise ValueError(f"wavespeed array must have the node shape {self.Nx}")
        return v.to(self.device)


from .acoustic import AcousticWave  # noqa: E402,F401  (cuwave.scalar exports both descriptors)
