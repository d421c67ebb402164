"""Uneven shot batch: the misfit gradients of many 2D shots of different grid sizes, orders and
step counts in one persistent kernel launch (``wfd_batch_*``), optionally replayed from a
captured hipGraph.

Each shot gets its own grid, medium, source, receivers and observed record.  One ``run`` performs
every shot's checkpointed forward run, residual and adjoint run; :meth:`ShotBatch.gradients`
turns the per-shot imaging sums into ``(J, dJ/dv)``.  Per shot, the result is bit-identical to the
single-shot :class:`~neuralmech_amd.solver.GradientPlan` path with the residual taken in the grid
dtype.  This replaces the per-shot loop of cuwave's ``misfit_gradient`` over sources
(``projects/18_fwi/2D_fwi.py:40``) for shots whose grids differ.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from ._lib import check
from .solver import Medium, PointSet, WaveGrid, _ptr, _sig_tensor, _stream_ptr, kc_gradient

_MODES = {"persistent": 0, "phases": 1}


@dataclass
class Shot:
    """One shot: node grid ``shape`` with ``spacing``, time step ``dt``, stencil ``order``,
    wave speed node array ``velocity``, damping ``beta_profiles`` (one per axis or None), source
    nodes and signal ``(nsteps,)`` / ``(nsteps, S)``, receiver nodes and the observed record
    ``(nsteps, R)``; ``seg`` is the checkpoint spacing (0 = memory-optimal)."""

    shape: Sequence[int]
    spacing: Sequence[float]
    dt: float
    order: int
    velocity: object
    src_nodes: Sequence
    signal: object
    rec_nodes: Sequence
    observed: object
    nsteps: int
    beta_profiles: Optional[Sequence] = None
    seg: int = 0


class ShotBatch:
    """Device state of a batch of :class:`Shot` (all of one dtype).

    ``max_active_cells`` bounds the interior cells of the shots in flight (0 = library default),
    ``mode`` is ``"persistent"`` (one launch) or ``"phases"`` (one launch per phase, same tiles).
    """

    def __init__(self, shots: Sequence[Shot], dtype=torch.float32, device=None, max_active_cells: int = 0,
                 blocks_per_cu: int = 0, mode: str = "persistent", timeout_ms: int = 0):
        if not shots:
            raise ValueError("empty batch")
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self.dtype = dtype
        self.shots = list(shots)
        self._keep = []
        arr = (_lib.BatchShot * len(shots))()
        self.grids, self.media, self.srcs, self.recs, self.vel = [], [], [], [], []
        self.sig, self.traces, self.obs, self.adj, self.mu_src, self.G = [], [], [], [], [], []
        for k, sh in enumerate(shots):
            g = WaveGrid(sh.shape, sh.spacing, sh.dt, sh.order, dtype, nshots=1, device=self.device)
            v64 = torch.as_tensor(sh.velocity, dtype=torch.float64)
            med = Medium(g, velocity=v64, beta_profiles=sh.beta_profiles)
            v64 = v64.to(self.device)  # kept on the device for dJ/dv = dJ/dkc 2 v dt^2
            src = PointSet(g, [tuple(int(c) for c in p) for p in sh.src_nodes])
            rec = PointSet(g, [tuple(int(c) for c in p) for p in sh.rec_nodes])
            N = int(sh.nsteps)
            sig = _sig_tensor(g, src, torch.as_tensor(sh.signal), N) if src.count else None
            obs = torch.as_tensor(sh.observed).to(self.device, dtype).reshape(1, N, rec.count).contiguous()
            traces = torch.zeros(1, N, rec.count, dtype=dtype, device=self.device)
            adj = torch.zeros_like(traces)
            mu = torch.zeros(1, N, max(src.count, 1), dtype=dtype, device=self.device)
            G = torch.zeros(g.elems, dtype=torch.float64, device=self.device)
            c = arr[k]
            c.grid = g.c
            c.nsteps = N
            c.seg = int(sh.seg)
            c.med = med.c
            sp = src.struct(sig) if sig is not None else None
            if sp is not None:
                c.src = sp
            rp = rec.struct(traces)
            if rp is not None:
                c.rec = rp
            c.observed = obs.data_ptr()
            c.adj = adj.data_ptr()
            c.mu_src = mu.data_ptr()
            c.G = G.data_ptr()
            self.grids.append(g)
            self.media.append(med)
            self.vel.append(v64)
            self.srcs.append(src)
            self.recs.append(rec)
            self.sig.append(sig)
            self.traces.append(traces)
            self.obs.append(obs)
            self.adj.append(adj)
            self.mu_src.append(mu)
            self.G.append(G)
        opts = _lib.BatchOpts()
        opts.max_active_cells = int(max_active_cells)
        opts.blocks_per_cu = int(blocks_per_cu)
        opts.mode = _MODES[mode]
        opts.timeout_ms = int(timeout_ms)
        self._p = ctypes.c_void_p()
        check(_lib.load().wfd_batch_create(ctypes.byref(self._p), len(shots), arr, ctypes.byref(opts)))
        self._arr = arr

    def __del__(self):
        p = getattr(self, "_p", None)
        if p is not None and p.value:
            try:
                _lib.load().wfd_batch_destroy(p)
            except Exception:
                pass
            self._p = None

    # ---------------------------------------------------------------------------- control
    def tune(self, steps: int = 20) -> list:
        """Autotune the tile shape of every size class; returns the per-shot (variant, rows)."""
        check(_lib.load().wfd_batch_tune(self._p, int(steps), _stream_ptr(self.device)))
        return [self.shot_info(k)[1:3] for k in range(len(self.shots))]

    def set_config(self, variant: int, rows: int, size_class: int = -1) -> None:
        check(_lib.load().wfd_batch_set_config(self._p, int(size_class), int(variant), int(rows)))

    def run(self, graph: bool = True) -> None:
        """Every shot's forward, residual and adjoint run (stream-ordered; no host sync)."""
        check(_lib.load().wfd_batch_run(self._p, 1 if graph else 0, _stream_ptr(self.device)))

    def status(self) -> None:
        """Synchronise and raise if the last run did not complete."""
        check(_lib.load().wfd_batch_status(self._p, _stream_ptr(self.device)))

    def info(self) -> dict:
        ph, bl, cs = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        by = ctypes.c_size_t()
        check(_lib.load().wfd_batch_info(self._p, ctypes.byref(ph), ctypes.byref(bl), ctypes.byref(by),
                                         ctypes.byref(cs)))
        return {"phases": ph.value, "blocks": bl.value, "bytes": by.value, "cell_steps": cs.value}

    def shot_info(self, k: int):
        """``(size_class, variant, rows, seg, start_phase, phases)`` of shot ``k``."""
        c, v, r = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        sg, st, ph = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(_lib.load().wfd_batch_shot_info(self._p, int(k), ctypes.byref(c), ctypes.byref(v), ctypes.byref(r),
                                              ctypes.byref(sg), ctypes.byref(st), ctypes.byref(ph)))
        return c.value, v.value, r.value, sg.value, st.value, ph.value

    # ---------------------------------------------------------------------------- results
    def imaging_sum(self, k: int) -> torch.Tensor:
        """``G = sum_n mu^{n+1} L u^n`` of shot ``k`` on its nodes (float64, before the source term)."""
        return self.grids[k].nodes(self.G[k][None])[0]

    def gradients(self):
        """Per shot ``(J, dJ/dv)``: ``J = 1/2 sum (d - d_obs)^2`` and the float64 gradient node array."""
        out = []
        for k in range(len(self.shots)):
            g = self.grids[k]
            r = self.traces[k].double() - self.obs[k].double()
            J = 0.5 * float((r * r).sum())
            src = self.srcs[k]
            mu = self.mu_src[k][:, :, : src.count] if src.count else None
            kc = self.media[k].kc_nodes()[0]
            gk = kc_gradient(g, self.imaging_sum(k).clone(), mu, self.sig[k], src if src.count else None, kc)
            out.append((J, gk * 2.0 * self.vel[k] * g.dt ** 2))
        return out


def shot_ops(nsteps: int, seg: int = 0, has_src: bool = True, has_rec: bool = True):
    """The op list the batch runs for one shot, ``(ops (n, 11) int32, nfields)`` (host only):
    columns kind, flags, up, uc, un, uf, gp, gf, step, row, phase (see ``wfd_batch_shot_ops``)."""
    lib = _lib.load()
    cnt, nf = ctypes.c_int64(), ctypes.c_int32()
    check(lib.wfd_batch_shot_ops(int(nsteps), int(seg), int(has_src), int(has_rec), None, 0, ctypes.byref(cnt),
                                 ctypes.byref(nf)))
    ops = np.zeros((cnt.value, 11), np.int32)
    check(lib.wfd_batch_shot_ops(int(nsteps), int(seg), int(has_src), int(has_rec),
                                 ops.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), cnt.value, ctypes.byref(cnt),
                                 ctypes.byref(nf)))
    return ops, nf.value
