"""``stable_dt``, ``grid_coords`` and ``simulate``: the ``cuwave.wave`` calls of the drivers
(``projects/18_fwi/helper.py:52``, ``plate.py:47,56``, ``8_physics_drivers/waveND.py:94-95``,
``18_fwi/animate_bhai_waves.py:29``) over the wavefd solver."""

from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from . import solver


def stable_dt(dx, wavespeed, space_order: int = 2) -> float:
    """Largest stable step for spacing ``dx`` (tuple), the largest of ``wavespeed`` and the order."""
    h = tuple(float(x) for x in (dx if np.ndim(dx) else [dx]))
    vmax = float(torch.as_tensor(wavespeed, dtype=torch.float64).max())
    return solver.stable_dt(h, vmax, int(space_order))


def grid_coords(Nx, dx, dtype=np.float64):
    """Node coordinates, one ``ij``-indexed array per axis: node ``i`` at ``(i - 1) dx``."""
    axes = [(np.arange(n, dtype=np.float64) - 1.0) * h for n, h in zip(Nx, dx if np.ndim(dx) else [dx] * len(Nx))]
    return [a.astype(dtype) for a in np.meshgrid(*axes, indexing="ij")]


def _unit_indicator(indicator) -> bool:
    """True when there is no density indicator (None): the constant-density kernels run.  Any
    indicator array, ones included, selects the variable-density form, so that runs with and
    without voids share one discretisation (``projects/18_fwi/2D_tfm.py:29-30`` subtracts them)."""
    return indicator is None


def _setup(sim, sources, sensors, nshots):
    g = solver.WaveGrid(sim.Nx, sim.dx, sim.dt, sim.space_order, sim.dtype, nshots=nshots, device=sim.device)
    med = solver.Medium(g, velocity=sim.velocity(), beta_profiles=sim.damping)
    if sources:
        nodes = [[tuple(int(c) for c in p) for p in s.idx.T] for s in sources]
        src = solver.PointSet(g, nodes, per_shot=nshots > 1) if nshots > 1 else solver.PointSet(g, nodes[0])
        width = max(s.signal.shape[1] for s in sources)
        sig = np.zeros((nshots, sim.N, width))
        for k, s in enumerate(sources):
            sig[k, :, : s.signal.shape[1]] = s.signal[: sim.N]
        sig = torch.as_tensor(sig)
    else:
        src, sig = None, None
    rec = solver.PointSet(g, [tuple(int(c) for c in p) for p in np.asarray(sensors.nodes).T]) \
        if sensors is not None else None
    return g, med, src, sig, rec


def simulate(sim, source=None, indicator=None, sensors=None, record_every: Optional[int] = None):
    """Run ``sim.N`` steps from rest.

    A density ``indicator`` (or an ``AcousticWave``) runs the variable mass / stiffness model.  Returns ``u`` (final field, node array on the device), ``(u, traces)`` with ``sensors``
    (``traces`` of shape ``(N, R)``), ``(u, snaps)`` with ``record_every`` (host arrays every that many
    steps), or ``(u, traces, frames)`` with both.
    """
    from .acoustic import AcousticWave, simulate_acoustic

    if isinstance(sim, AcousticWave) or not _unit_indicator(indicator):
        return simulate_acoustic(sim, source, indicator, sensors, record_every)
    sources = [source] if source is not None else []
    g, med, src, sig, rec = _setup(sim, sources, sensors, 1)
    fr = solver.ForwardRun(g, med, sim.N, src=src, sig=sig, rec=rec)
    frames = []
    chunk = int(record_every) if record_every else sim.N
    while fr.state.step < sim.N:
        fr.advance(min(chunk, sim.N - fr.state.step))
        if record_every:
            frames.append(g.nodes(fr.state.cur)[0].detach().cpu().numpy().copy())
    u = g.nodes(fr.state.cur)[0].clone()
    out = [u]
    if sensors is not None:
        out.append(fr.traces[0] if fr.traces is not None
                   else torch.zeros(sim.N, 0, dtype=sim.dtype, device=sim.device))
    if record_every:
        out.append(frames)
    return out[0] if len(out) == 1 else tuple(out)
