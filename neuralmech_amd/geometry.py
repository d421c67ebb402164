"""Node-set helpers of the call sites: ``box`` and ``nodes``
(``projects/8_physics_drivers/topopt_acoustic2D.py:144,149-153``), ``circle`` (the voids of
``projects/18_fwi/plate.py:54-61`` are stacked circles) and ``random_ellipses``
(``projects/16_neuralopt/helper.py:72-79``)."""

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


def random_ellipses(coords, count: int, radii, bounds, overlap: bool = True, rng=None, max_tries: int = 10000):
    """Mask of ``count`` random elliptical voids (2D), as ``projects/16_neuralopt/helper.py:72-79``
    asks for them: semi-axes uniform in ``radii = (r_min, r_max)``, orientation uniform in
    ``[0, pi)``, centres uniform in ``bounds = [(x_lo, x_hi), (y_lo, y_hi)]``.  With
    ``overlap=False`` a candidate is redrawn while its circumscribed circle (radius = the larger
    semi-axis) meets that of an accepted ellipse, so the voids are disjoint; ``RuntimeError`` after
    ``max_tries`` draws.  cuwave's sampling is not available; this is this build's reading of the
    call site's arguments.  ``rng`` is a ``numpy.random.Generator`` (a fresh default one if None)."""
    rng = np.random.default_rng() if rng is None else rng
    c = [np.asarray(a.get() if hasattr(a, "get") else a, dtype=np.float64) for a in coords]
    if len(c) != 2 or len(bounds) != 2:
        raise ValueError("random_ellipses: 2D coordinates and bounds")
    r_lo, r_hi = (float(v) for v in radii)
    if not 0 < r_lo <= r_hi:
        raise ValueError("random_ellipses: 0 < r_min <= r_max")
    accepted = []
    tries = 0
    while len(accepted) < int(count):
        tries += 1
        if tries > max_tries:
            raise RuntimeError(f"random_ellipses: placed {len(accepted)} of {count} disjoint ellipses")
        a, b = rng.uniform(r_lo, r_hi, 2)
        theta = rng.uniform(0.0, np.pi)
        centre = np.array([rng.uniform(*bounds[0]), rng.uniform(*bounds[1])])
        rad = max(a, b)
        if not overlap and any(np.hypot(*(centre - q)) <= rad + rq for q, rq, *_ in accepted):
            continue
        accepted.append((centre, rad, a, b, theta))
    mask = np.zeros(c[0].shape, bool)
    for centre, _, a, b, theta in accepted:
        dx, dy = c[0] - centre[0], c[1] - centre[1]
        u = np.cos(theta) * dx + np.sin(theta) * dy
        v = -np.sin(theta) * dx + np.cos(theta) * dy
        mask |= (u / a) ** 2 + (v / b) ** 2 <= 1.0
    return mask
