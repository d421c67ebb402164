"""Differentiable forward modelling: ``torch.autograd.Function`` around the checkpointed plan.

``traces = wave_traces(v, grid, src, sig, rec, ...)`` runs the HIP forward from rest and keeps
only sqrt(n)-spaced checkpoints; ``backward`` recomputes each segment and runs the hand-written
adjoint kernel (exact discrete transpose of the step) with ``-dL/dtraces`` injected at the
receivers, returning ``dL/dv``.  This is the drop-in for cuwave's ``misfit_gradient``
(``projects/18_fwi/2D_fwi.py:40``) in the velocity parametrisation.
"""

from __future__ import annotations

import ctypes
from typing import Optional

import torch

from .solver import GradientPlan, Medium, PointSet, WaveGrid

# The last gradient plan, kept for the next call on an identical grid descriptor: an FWI loop
# reuses its checkpoint store (every forward restarts from rest and every backward clears its
# adjoint fields, so a plan runs any number of forward / backward pairs) instead of allocating and
# clearing it each iteration.  A plan is leased to one forward until its backward has run (or its
# graph is dropped); a call while the kept plan is leased gets a plan of its own.
# clear_plans() releases the kept plan.
_PLAN_CACHE: dict = {}


class _Lease:
    """Marks a plan as holding one forward's checkpoints until released (or collected)."""

    def __init__(self, plan):
        self.plan = plan
        plan._leased = True

    def release(self):
        if self.plan is not None:
            self.plan._leased = False
            self.plan = None

    def __del__(self):
        self.release()


def _plan_for(grid: WaveGrid, nsteps: int, seg: int, max_bytes: int) -> GradientPlan:
    key = (ctypes.string_at(ctypes.addressof(grid.c), ctypes.sizeof(grid.c)), str(grid.device), int(nsteps),
           int(seg), int(max_bytes))
    plan = _PLAN_CACHE.get(key)
    if plan is not None and getattr(plan, "_leased", False):
        return GradientPlan(grid, nsteps, seg=seg, max_bytes=max_bytes)  # the kept one is in use
    if plan is None:
        _PLAN_CACHE.clear()  # one plan at a time: plans can hold tens of GB
        plan = GradientPlan(grid, nsteps, seg=seg, max_bytes=max_bytes)
        _PLAN_CACHE[key] = plan
    return plan


def clear_plans() -> None:
    """Release the gradient plan kept by :func:`wave_traces` / :func:`misfit_gradient`."""
    _PLAN_CACHE.clear()


class _WaveTraces(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v, grid: WaveGrid, src: Optional[PointSet], sig, rec: PointSet, beta_profiles, seg, max_bytes):
        vd = v.detach().to(torch.float64)
        med = Medium(grid, velocity=vd, beta_profiles=beta_profiles)
        plan = _plan_for(grid, _nsteps(sig), seg, max_bytes)
        ctx.lease = _Lease(plan)
        traces = plan.forward(med, src, sig, rec)
        ctx.plan, ctx.med, ctx.src, ctx.rec, ctx.grid = plan, med, src, rec, grid
        ctx.save_for_backward(vd)
        return traces

    @staticmethod
    def backward(ctx, grad_traces):
        (vd,) = ctx.saved_tensors
        g = ctx.grid
        grad_kc = ctx.plan.backward(ctx.med, ctx.src, ctx.rec, grad_traces.contiguous())
        grad_v = grad_kc * 2.0 * vd * g.dt ** 2
        ctx.plan = None
        ctx.lease.release()
        return grad_v.to(torch.float64), None, None, None, None, None, None, None


def _nsteps(sig) -> int:
    if sig is None:
        raise ValueError("a source signal fixes the number of steps")
    return int(sig.shape[-2]) if sig.dim() >= 2 else int(sig.shape[0])


def wave_traces(v: torch.Tensor, grid: WaveGrid, src: Optional[PointSet], sig: torch.Tensor, rec: PointSet,
                beta_profiles=None, seg: int = 0, max_bytes: int = 0) -> torch.Tensor:
    """Receiver traces ``(nshots, nsteps, R)`` for the velocity node array ``v`` (differentiable in
    ``v``; float64 gradient).  ``sig`` is ``(nsteps, S)`` or ``(nshots, nsteps, S)``."""
    return _WaveTraces.apply(v, grid, src, sig, rec, beta_profiles, seg, max_bytes)


def misfit_gradient(v, grid, src, sig, rec, observed, beta_profiles=None, seg=0, max_bytes=0):
    """``(J, dJ/dv)`` for ``J = 1/2 sum (d - d_obs)^2`` over all shots (one forward + one adjoint)."""
    vv = torch.as_tensor(v, dtype=torch.float64, device=grid.device).detach().requires_grad_(True)
    d = wave_traces(vv, grid, src, torch.as_tensor(sig, device=grid.device), rec, beta_profiles, seg, max_bytes)
    r = d.double() - torch.as_tensor(observed, device=grid.device).double()
    J = 0.5 * (r * r).sum()
    J.backward()
    return float(J.detach()), vv.grad.detach()
