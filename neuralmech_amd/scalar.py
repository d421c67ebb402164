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
* ``wavespeed`` is a number or a node array.  Without an indicator the model is
  ``u_tt = v^2 lap u`` and ``density`` does not enter; with a density indicator ``gamma`` (``simulate``,
  ``measure``, ``misfit``, ``misfit_gradient``) it is the variable-density form
  ``gamma rho u_tt = div(gamma rho v^2 grad u) + rho v^2 f`` on the zero boundary ring, run by the
  variable mass / stiffness kernels (2D).
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

    # ---- density indicator (the variable-density form the indicator of misfit_gradient scales)
    def _rho(self) -> torch.Tensor:
        r = torch.as_tensor(self.density, dtype=torch.float64).to(self.device)
        return r.expand(self.Nx) if r.dim() == 0 else r

    def _gamma(self, gamma) -> torch.Tensor:
        g = torch.as_tensor(gamma if gamma is not None else 1.0, dtype=torch.float64).to(self.device)
        g = g.expand(self.Nx) if g.dim() == 0 else g
        if tuple(g.shape) != self.Nx:
            raise ValueError(f"indicator must have the node shape {self.Nx}")
        return g

    def material(self, gamma):
        """``(mass, stiffness)`` of ``rho u_tt = div(rho c^2 grad u)`` with the density scaled by the
        indicator: ``(gamma rho, gamma rho c^2)`` (2D)."""
        g = self._gamma(gamma)
        rho = self._rho()
        return g * rho, g * rho * self.velocity() ** 2

    def parametrization_jacobian(self, gamma):
        """``(d mass / d gamma, d stiffness / d gamma) = (rho, rho c^2)``."""
        self._gamma(gamma)
        rho = self._rho()
        return rho.clone(), rho * self.velocity() ** 2

    def source_scale(self) -> torch.Tensor:
        """Sources enter as ``rho c^2 f`` so that a unit indicator reproduces ``u_tt = c^2 (lap u + f)``
        of the constant-density model where ``c`` is constant."""
        return self._rho() * self.velocity() ** 2

    def grid(self, nshots: int = 1):
        from .solver import WaveGrid

        return WaveGrid(self.Nx, self.dx, self.dt, self.space_order, self.dtype, nshots=nshots, device=self.device)

    def velocity(self) -> torch.Tensor:
        """Wave speed as a float64 node array on the device."""
        v = torch.as_tensor(self.wavespeed, dtype=torch.float64)
        if v.dim() == 0:
            v = v.expand(self.Nx)
        if tuple(v.shape) != self.Nx:
            raise ValueError(f"wavespeed array must have the node shape {self.Nx}")
        return v.to(self.device)


from .acoustic import AcousticWave  # noqa: E402,F401  (cuwave.scalar exports both descriptors)
