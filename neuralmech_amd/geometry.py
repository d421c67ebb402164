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


def stacked_circles(coords, num: int, max_radius: float, span, position: float, axis: int = 1,
                    ratio: float = 0.5, order: str = "ascending"):
    """Mask of ``num`` discs stacked along ``axis`` (``projects/18_fwi/plate.py:56-59``): centres
    evenly spaced over ``span`` = (start, end) (the middles of ``num`` equal segments) at
    ``position`` on the other axis, radii from ``ratio * max_radius`` to ``max_radius`` in
    ``order`` ("ascending" along the axis, or "descending").  cuwave's exact placement is not
    available; this is this build's reading of the call site's arguments."""
    import numpy as _np

    radii = max_radius * _np.linspace(ratio, 1.0, int(num))
    if order == "descending":
        radii = radii[::-1]
    lo, hi = (float(v) for v in span)
    centres = lo + (hi - lo) * (_np.arange(int(num)) + 0.5) / int(num)
    mask = None
    for c, r in zip(centres, radii):
        centre = [0.0, 0.0]
        centre[axis] = c
        centre[1 - axis] = float(position)
        m = circle(coords, centre, r)
        mask = m if mask is None else mask | m
    return mask
