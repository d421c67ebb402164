"""Node-set helpers of the call sites: ``box`` and ``nodes``
(``projects/8_physics_drivers/topopt_acoustic2D.py:144,149-153``) and ``circle`` (the voids of
``projects/18_fwi/plate.py:54-61`` are stacked circles)."""

from __future__ import annotations

import numpy as np


def box(coords, center, size):
    """Boolean node mask of the axis-aligned box ``center +- size / 2`` (closed) over the coordinate
    arrays ``coords`` (one per axis, as ``wave.grid_coords`` returns)."""
    c = [np.asarray(a.get() if hasattr(a, "get") else a, dtype=np.float64) for a in coords]
    mask = np.ones(c[0].shape, bool)
    for a, m, s in zip(c, center, size):
        mask &= np.abs(a - float(m)) <= 0.5 * float(s) + 1e-12 * max(1.0, abs(float(s)))
    return mask


def nodes(mask):
    """Node indices ``(D, S)`` of a boolean node mask, in C order."""
    return np.argwhere(np.asarray(mask, bool)).T.copy()


def circle(coords, center, radius):
    """Boolean node mask of the closed disc ``|x - center| <= radius`` (2D) over ``coords``."""
    c = [np.asarray(a.get() if hasattr(a, "get") else a, dtype=np.float64) for a in coords]
    r2 = sum((a - float(m)) ** 2 for a, m in zip(c, center))
    return r2 <= float(radius) ** 2 * (1 + 1e-12)
