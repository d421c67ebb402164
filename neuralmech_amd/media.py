"""Seeded synthetic velocity models and acquisition geometry for the benchmark workloads.

All randomness comes from ``numpy.random.default_rng(seed)``; every model is a float64 node array
including the boundary ring.
"""

from __future__ import annotations

import math

import numpy as np
from scipy.ndimage import gaussian_filter


def smoothed_random(shape, seed, sigma, vmin, vmax) -> np.ndarray:
    """White noise, Gaussian-filtered with ``sigma`` cells, min-max mapped to ``[vmin, vmax]``."""
    rng = np.random.default_rng(seed)
    p = gaussian_filter(rng.standard_normal(shape), sigma, mode="reflect")
    p = (p - p.min()) / (p.max() - p.min())
    return vmin + (vmax - vmin) * p


def forward_2d_model(n=2048, seed=2) -> np.ndarray:
    """2D single-shot model: smoothed random field (sigma 20) on [1500, 4500] m/s plus three
    horizontal interfaces at depths n/4, n/2, 3n/4 each adding 300 m/s, clipped to [1500, 4500]."""
    v = smoothed_random((n, n), seed, 20.0, 1500.0, 4500.0)
    for depth in (n // 4, n // 2, 3 * n // 4):
        v[depth:, :] += 300.0
    return np.clip(v, 1500.0, 4500.0)


def layered_fwi_models(n=1024, seed=3):
    """FWI pair: true = 4 layers 1500 -> 4500 m/s times (1 + 0.1 p), p smoothed noise in [-1, 1];
    start = true filtered with sigma 40 cells."""
    base = np.empty((n, n))
    speeds = np.linspace(1500.0, 4500.0, 4)
    edges = np.linspace(0, n, 5).astype(int)
    for k in range(4):
        base[edges[k]:edges[k + 1], :] = speeds[k]
    p = smoothed_random((n, n), seed, 20.0, -1.0, 1.0)
    true = base * (1.0 + 0.1 * p)
    start = gaussian_filter(true, 40.0, mode="nearest")
    return true, start


def volume_3d_model(n=512, seed=5, device=None):
    """3D model: white noise from ``default_rng(seed)``, Gaussian-smoothed with sigma 20 cells
    (periodic, by FFT on the GPU when one is given, so a 512^3 model takes seconds), min-max
    mapped to [1500, 4500] m/s.  Returns a float64 torch tensor on ``device`` (CPU if None)."""
    import torch

    rng = np.random.default_rng(seed)
    dev = torch.device(device) if device is not None else torch.device("cpu")
    noise = torch.from_numpy(rng.standard_normal((n, n, n))).to(dev)
    k = [torch.fft.fftfreq(n, device=dev, dtype=torch.float64) for _ in range(2)] + [
        torch.fft.rfftfreq(n, device=dev, dtype=torch.float64)]
    sigma = 20.0
    spec = torch.fft.rfftn(noise)
    del noise
    for axis, kk in enumerate(k):
        shape = [1, 1, 1]
        shape[axis] = kk.numel()
        spec *= torch.exp(-2.0 * (math.pi * sigma * kk) ** 2).reshape(shape)
    p = torch.fft.irfftn(spec, s=(n, n, n))
    del spec
    lo, hi = p.min(), p.max()
    return 1500.0 + 3000.0 * (p - lo) / (hi - lo)


def ricker(t, f0: float, t0: float | None = None) -> np.ndarray:
    """Ricker wavelet ``(1 - 2 pi^2 f0^2 tau^2) exp(-pi^2 f0^2 tau^2)``, ``tau = t - t0``, ``t0 = 1/f0``."""
    t0 = 1.0 / f0 if t0 is None else t0
    a = (math.pi * f0 * (np.asarray(t, dtype=np.float64) - t0)) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


def sponge_profile(n: int, width: int, beta_max: float, lo: bool = True, hi: bool = True) -> np.ndarray:
    """Quadratic damping ``beta(i) = beta_max ((W - e)/W)^2`` for edge distance ``e < W`` (nodes from the
    boundary ring), on the faces flagged ``lo`` / ``hi``; ``beta = d dt / 2`` per step."""
    i = np.arange(n, dtype=np.float64)
    beta = np.zeros(n)
    if width > 0:
        for on, e in ((lo, i), (hi, n - 1 - i)):
            if on:
                beta = np.maximum(beta, np.where(e < width, beta_max * ((width - e) / width) ** 2, 0.0))
    return beta
