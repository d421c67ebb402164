"""Source wavelets: ``sineburst`` (the wavelet of ``projects/18_fwi/helper.py:74`` and
``plate.py:51``) and ``ricker`` (the configs' wavelet)."""

from __future__ import annotations

import numpy as np

from .media import ricker  # noqa: F401  (re-exported)


def sineburst(t, amplitude: float, frequency: float, cycles: float) -> np.ndarray:
    """``A sin(2 pi f t)`` under a Hann envelope ``(1 - cos(2 pi f t / cycles)) / 2`` for
    ``0 <= t <= cycles / f``, zero afterwards (this build's definition; cuwave's
    ``cuwave.signals.sineburst`` is not available to compare against)."""
    t = np.asarray(t, dtype=np.float64)
    on = (t >= 0) & (t <= cycles / frequency)
    env = 0.5 * (1.0 - np.cos(2.0 * np.pi * frequency * t / cycles))
    return np.where(on, amplitude * np.sin(2.0 * np.pi * frequency * t) * env, 0.0)
