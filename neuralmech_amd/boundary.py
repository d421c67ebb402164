"""Absorbing layers: ``sponge`` and ``pad_for_sponge``, the calls of
``projects/8_physics_drivers/topopt_acoustic2D.py:104,132-134``.

The damping is the damped leapfrog of the solver, ``(1+b) u+ = 2u - (1-b) u- + kc (L u + s)``,
with ``b`` per node the sum of one quadratic profile per axis, ``b_max ((W - e) / W)^2`` over the
``W`` nodes nearest each damped face (``e`` counted from the boundary ring).  ``beta`` is the
peak ``d dt / 2`` at the wall, the parametrisation of ``topopt_acoustic2D.py:68-69``.
"""

from __future__ import annotations

import math

import numpy as np

from .media import sponge_profile


def _faces(ndim, faces):
    """Set of (axis, side) with side 0 = low face, 1 = high face."""
    if faces is None or faces == "all":
        return {(d, s) for d in range(ndim) for s in (0, 1)}
    out = set()
    for f in faces:
        d, s = f
        s = {"lo": 0, "low": 0, "-": 0, -1: 0, 0: 0, "hi": 1, "high": 1, "+": 1, 1: 1}[s]
        if not 0 <= d < ndim:
            raise ValueError(f"face axis {d} out of range")
        out.add((int(d), s))
    return out


def sponge(sim, material=None, width=20, beta=0.05, faces="all"):
    """Damping profiles for ``dataclasses.replace(sim, damping=...)``: one array per axis.

    ``width`` is the layer thickness in nodes, ``beta`` the peak damping per step at the wall;
    ``material`` (the medium filling the layer at the call sites) does not change the profile.
    """
    fs = _faces(sim.ndim, faces)
    return tuple(sponge_profile(n, int(width), float(beta), lo=(d, 0) in fs, hi=(d, 1) in fs)
                 for d, n in enumerate(sim.Nx))


def pad_for_sponge(resolution, dx, thickness, faces="all"):
    """Grow a region by a sponge of ``thickness`` (length units) on the chosen faces.

    Returns ``(Nx, width, origin, region)``: the padded node counts, the layer width in nodes,
    the coordinate of the padded grid's node 1 relative to the region's node 1 (negative on padded
    low faces), and the tuple of slices selecting the region's nodes in padded node arrays.
    """
    res = [int(r) for r in np.atleast_1d(resolution)]
    dxs = [float(h) for h in (dx if np.ndim(dx) else [dx] * len(res))]
    width = int(math.ceil(float(thickness) / min(dxs) - 1e-12))
    fs = _faces(len(res), faces)
    Nx, origin, region = [], [], []
    for d, (n, h) in enumerate(zip(res, dxs)):
        lo = width if (d, 0) in fs else 0
        hi = width if (d, 1) in fs else 0
        Nx.append(n + lo + hi)
        origin.append(-lo * h)
        region.append(slice(lo, lo + n))
    return tuple(Nx), width, tuple(origin), tuple(region)
