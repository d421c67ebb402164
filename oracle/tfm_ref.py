"""CPU restatement of the total focusing method of ``projects/18_fwi/tfm.py``.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this module; ``tests/`` and the
``cpu_baseline`` leg of ``bench.py --workload tfm`` use it.

* :func:`tfm_all_pairs` is ``tfm`` of ``tfm.py:39-53`` (array operations over all E^2 pairs,
  ``k = min(int(t), Nt - 2)``, the sample kept where ``k < Nt - 2``), in NumPy.
* :func:`tfm_reciprocal` is the RawKernel of ``tfm.py:12-36`` written with NumPy arrays: pairs
  ``r >= e``, the pair ``r > e`` weighted 2, samples with ``k >= Nt - 1`` skipped.
* :func:`hilbert` is ``scipy.signal.hilbert`` (the library routine ``cupyx.scipy.signal.hilbert``
  mirrors, ``2D_tfm.py:35``), the pin of the package's FFT analytic signal.

Pinned by tests/test_tfm.py: the image of a point scatterer's synthetic capture peaks at the
scatterer; with reciprocal traces the two forms agree away from the end of the record; a pixel
summed by an explicit loop over pairs equals the array form.
"""

from __future__ import annotations

import numpy as np


def _delays(elements, x, y, c):
    return [np.hypot(x - ex, y - ey) / c for ex, ey in np.asarray(elements, np.float64)]


def tfm_all_pairs(h, elements, t0, dt, c, x, y):
    h = np.asarray(h)
    E, _, Nt = h.shape
    tau = _delays(elements, x, y, c)
    s = np.zeros(np.shape(x), np.complex128)
    for e in range(E):
        for r in range(E):
            t = (t0 + tau[e] + tau[r]) / dt
            k = np.minimum(t.astype(np.int32), Nt - 2)
            w = t - k
            s += ((1 - w) * h[e, r, k] + w * h[e, r, k + 1]) * (k < Nt - 2)
    return np.abs(s)


def tfm_reciprocal(h, elements, t0, dt, c, x, y):
    h = np.asarray(h)
    E, _, Nt = h.shape
    tau = _delays(elements, x, y, c)
    s = np.zeros(np.shape(x), np.complex128)
    for e in range(E):
        for r in range(e, E):
            t = (t0 + tau[e] + tau[r]) / dt
            k = t.astype(np.int64)
            ok = k < Nt - 1
            kk = np.where(ok, k, 0)
            f = 2.0 if r > e else 1.0
            w = (t - k) * f
            v = f - w
            s += np.where(ok, v * h[e, r, kk] + w * h[e, r, kk + 1], 0)
    return np.abs(s)


def hilbert(x, axis=-1):
    from scipy.signal import hilbert as _h

    return _h(x, axis=axis)
