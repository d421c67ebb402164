"""Sources, sensors and FWI helpers: the ``cuwave.utils`` calls of the drivers
(``projects/18_fwi/helper.py:74``, ``plate.py:42,51,62-64``, ``2D_fwi.py:31,40,47``,
``9_structural_optimization/fwi_penalization.py:99-143``, ``16_neuralopt/neuralfwi.py:60-62``)
over the wavefd solver.

Conventions of this build: points snap to the nearest node (node ``i`` at ``(i - 1) dx``,
clipped to the interior); source signals are divided by the cell volume, as cuwave's
``point_source`` does (``9_structural_optimization/README.md:87-88``); the model argument of
``misfit`` / ``misfit_gradient`` is the density indicator (variable-density model, see
``scalar.ScalarWave.material``); ``wavespeed_misfit(_gradient)`` take the wave speed instead.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Sequence

import numpy as np
import torch

from . import solver
from .autograd import misfit_gradient as _misfit_gradient
from .acoustic import energy, response, response_gradient  # noqa: F401  (cuwave.utils names)
from .wave import _setup, _unit_indicator


@dataclass
class Source:
    """Point source: node indices ``idx`` ``(D, S)`` and signal ``(N, S)`` (already / cell volume)."""

    idx: np.ndarray
    signal: np.ndarray

    def __post_init__(self):
        self.idx = np.asarray(self.idx, dtype=np.int32).reshape(len(self.idx), -1)
        sig = np.asarray(self.signal, dtype=np.float64)
        self.signal = sig[:, None] if sig.ndim == 1 else sig
        if self.signal.shape[1] != self.idx.shape[1]:
            raise ValueError("signal needs one column per source node")


def _snap(sim, coords) -> np.ndarray:
    """Nearest interior node of each coordinate, ``(D, S)``."""
    c = np.atleast_2d(np.asarray(coords, dtype=np.float64))
    if c.shape[-1] != sim.ndim:
        c = c.T
    idx = np.rint(c / np.asarray(sim.dx)) + 1
    hi = np.asarray(sim.Nx) - 2
    return np.clip(idx, 1, hi).astype(np.int32).T


def point_source(sim, coords, signal) -> Source:
    """Source(s) at the nearest node(s); ``signal`` ``(N,)`` or ``(N, S)`` divided by the cell volume."""
    idx = _snap(sim, coords)
    sig = np.asarray(signal, dtype=np.float64)
    if sig.ndim == 1:
        sig = np.repeat(sig[:, None], idx.shape[1], axis=1)
    return Source(idx, sig / sim.cell_volume)


def shots(sim, coords, signal) -> list:
    """One single-node ``Source`` per coordinate (full matrix capture: every element fires once)."""
    return [point_source(sim, c, signal) for c in np.atleast_2d(np.asarray(coords, dtype=np.float64))]


def line(p0, p1, count: int) -> np.ndarray:
    """``count`` evenly spaced points from ``p0`` to ``p1``, ``(count, D)``."""
    p0, p1 = np.asarray(p0, np.float64), np.asarray(p1, np.float64)
    t = np.linspace(0.0, 1.0, int(count))[:, None]
    return p0 + t * (p1 - p0)


class Sensors:
    """Receivers at the nodes nearest ``coords``; ``nodes`` is ``(D, R)`` node indices, the shape
    the call sites read (``topopt_acoustic2D.py:144,154`` prints ``sensors.shape[1]`` nodes;
    ``analog_rnn_train.py:143`` passes ``sensors.nodes`` on)."""

    def __init__(self, sim, coords):
        self.nodes = _snap(sim, coords)

    @property
    def count(self) -> int:
        return self.nodes.shape[1]

    def traces(self, record):
        """A recorded ``(N, R)`` trace array (identity: the solver already samples the nodes)."""
        return record

    def objective(self, fn):
        """``fn(traces (N, R)) -> (cost, dcost/dtraces)`` as an objective on the traces at
        ``self.nodes`` (``analog_rnn_train.py:143``)."""

        def objective(traces):
            return fn(self.traces(traces))

        return objective


def measure(sim, sources: Sequence[Source], indicator, sensors: Sensors) -> list:
    """Traces ``(N, R)`` of every source, all shots batched in one launch per time step.  A density
    indicator (or an ``AcousticWave``) runs the variable mass / stiffness model."""
    from .acoustic import AcousticWave, measure_indicator

    if isinstance(sim, AcousticWave) or not _unit_indicator(indicator):
        return measure_indicator(sim, list(sources), indicator, sensors)
    g, med, src, sig, rec = _setup(sim, list(sources), sensors, len(sources))
    tr, _ = solver.run(g, med, sim.N, src=src, sig=sig, rec=rec)
    return [tr[k] for k in range(len(sources))]


def _observed(observed, sim, nshots):
    obs = torch.stack([torch.as_tensor(o) for o in observed]) if isinstance(observed, (list, tuple)) \
        else torch.as_tensor(observed)
    if obs.dim() == 2:
        obs = obs[None]
    if obs.shape[0] != nshots:
        raise ValueError("one observed record per source")
    return obs.to(sim.device, torch.float64)


def misfit(sim, sources, gamma, sensors, observed) -> float:
    """``1/2 sum (d - d_obs)^2`` over all shots for the density indicator ``gamma``
    (``9_structural_optimization/fwi_penalization.py:139``; ``16_neuralopt/neuralfwi.py:60``)."""
    from .acoustic import shots_misfit

    return shots_misfit(sim, sources, gamma, sensors, _observed(observed, sim, len(sources)), gradient=False)


def misfit_gradient(sim, sources, gamma, sensors, observed):
    """``(cost, dcost/dgamma)`` on the nodes for the density indicator ``gamma``
    (``projects/18_fwi/2D_fwi.py:40``): one batched forward recording the reconstruction strip and
    one hand-written adjoint sweep over all shots (zero gradient where a sponge damps)."""
    from .acoustic import shots_misfit

    return shots_misfit(sim, sources, gamma, sensors, _observed(observed, sim, len(sources)), gradient=True)


def wavespeed_misfit(sim, sources, model, sensors, observed) -> float:
    """``1/2 sum (d - d_obs)^2`` over all shots for the wave speed node array ``model``."""
    from dataclasses import replace

    s2 = replace(sim, wavespeed=torch.as_tensor(model, dtype=torch.float64))
    d = torch.stack(measure(s2, sources, None, sensors)).double()
    r = d - _observed(observed, sim, len(sources))
    return 0.5 * float((r * r).sum())


def wavespeed_misfit_gradient(sim, sources, model, sensors, observed):
    """``(cost, dcost/dv)`` in the wave speed parametrisation: one batched forward with sqrt(n)
    checkpoints and one hand-written adjoint pass over all shots."""
    g, _, src, sig, rec = _setup(sim, list(sources), sensors, len(sources))
    v = torch.as_tensor(model, dtype=torch.float64, device=sim.device)
    return _misfit_gradient(v, g, src, sig, rec, _observed(observed, sim, len(sources)), sim.damping)


def interior_slice(sim) -> tuple:
    """Slices selecting the interior nodes (the boundary ring removed)."""
    return tuple(slice(1, n - 1) for n in sim.Nx)


def resample(record, dt_from: float, dt_to: float, N: int):
    """Linear interpolation of an ``(M, R)`` record sampled every ``dt_from`` onto ``N`` samples every
    ``dt_to`` (both starting at t = 0; beyond the record the last sample is held)."""
    is_t = torch.is_tensor(record)
    r = record.detach().cpu().double().numpy() if is_t else np.asarray(record, np.float64)
    t_from = np.arange(r.shape[0]) * dt_from
    t_to = np.arange(int(N)) * dt_to
    out = np.stack([np.interp(t_to, t_from, r[:, k]) for k in range(r.shape[1])], axis=1) if r.ndim == 2 \
        else np.interp(t_to, t_from, r)
    if is_t:
        return torch.as_tensor(out, dtype=record.dtype, device=record.device)
    return out


def threshold(gamma, t: float = 0.5):
    """Binary indicator: 1 where ``gamma >= t``, 0 elsewhere (``topopt_acoustic2D.py:258``)."""
    g = torch.as_tensor(gamma.get() if hasattr(gamma, "get") else gamma)
    return (g >= t).to(g.dtype if g.is_floating_point() else torch.float64)
