"""The batched FWI gradient workload: 1024 x 1024 nodes, 8th order, layered model with a seeded
+-10 % perturbation, smoothed starting model, 16 shots on the surface, 201 receivers, 3000 steps,
L2 trace misfit, gradient w.r.t. the velocity.

``batched_gradient`` runs every shot in one batched launch per step through the checkpointed
HIP forward + adjoint; ``plain_gradient`` is the per-shot baseline: the slice-form step written in
torch ops on the GPU, recorded by autograd, then ``loss.backward()``.
"""

from __future__ import annotations

import math
import time

import numpy as np
import torch
import torch.nn.functional as F

from . import media, solver
from .autograd import misfit_gradient

SPACING = (10.0, 10.0)


def setup(n=1024, nshots=16, nrec=201, nsteps=3000, dtype=torch.float32, seed=3):
    true, start = media.layered_fwi_models(n, seed)
    dt = 0.9 * solver.stable_dt(SPACING, float(max(true.max(), start.max())), 8)
    cols = np.round(np.linspace(64, n - 65, nshots)).astype(int)
    shots = [[(45, int(c))] for c in cols]
    rec = [(45, int(c)) for c in np.round(np.linspace(41, n - 42, nrec)).astype(int)]
    sig = media.ricker(np.arange(nsteps) * dt, 10.0) / (SPACING[0] * SPACING[1])
    beta = [media.sponge_profile(n, 40, 0.05) for _ in range(2)]
    g = solver.WaveGrid((n, n), SPACING, dt, 8, dtype, nshots=nshots)
    src = solver.PointSet(g, shots, per_shot=True)
    recs = solver.PointSet(g, rec)
    sigt = torch.as_tensor(np.broadcast_to(sig[None, :, None], (nshots, nsteps, 1)).copy())
    return dict(grid=g, src=src, rec=recs, sig=sigt, beta=beta, true=true, start=start, dt=dt, n=n,
                shots=shots, rec_nodes=rec, nsteps=nsteps, nshots=nshots)


def observed(w):
    """Traces of the true model (the 'measured' data)."""
    g = w["grid"]
    med = solver.Medium(g, velocity=torch.as_tensor(w["true"]), beta_profiles=w["beta"])
    tr, _ = solver.run(g, med, w["nsteps"], src=w["src"], sig=w["sig"], rec=w["rec"])
    return tr


def batched_gradient(w, obs, v=None, seg=0):
    """J and dJ/dv of all shots with sqrt(n) checkpointing (seg 0); ``seg=solver.SEG_FASTEST``
    keeps every forward level when it fits (no recompute in the backward pass: 20 + 22 bytes per
    cell-step instead of 16 + 20 + 22), with the same gradient bits."""
    v = w["start"] if v is None else v
    return misfit_gradient(torch.as_tensor(v), w["grid"], w["src"], w["sig"], w["rec"], obs, w["beta"], seg=seg)


def central_weights(order: int) -> list:
    """``a_0 .. a_K`` of the central second difference of ``order`` (closed form; SURVEY
    Appendix B's table: order 8 gives -205/72, 8/5, -1/5, 8/315, -1/560)."""
    K = order // 2
    a = [2.0 * (-1) ** (k + 1) * math.factorial(K) ** 2 / (k * k * math.factorial(K - k) * math.factorial(K + k))
         for k in range(1, K + 1)]
    return [-2.0 * sum(a)] + a


def lap_slices(u, h, order: int = 8):
    """The slice-form Laplacian of ``order`` with zero extension (torch ops, differentiable)."""
    a = central_weights(order)
    K = order // 2
    p = F.pad(u, (K, K, K, K))
    n0, n1 = u.shape
    c = p[K:K + n0, K:K + n1]
    lap = (a[0] / h[0] ** 2 + a[0] / h[1] ** 2) * c
    for k in range(1, K + 1):
        lap = lap + (a[k] / h[0] ** 2) * (p[K - k:K - k + n0, K:K + n1] + p[K + k:K + k + n0, K:K + n1])
        lap = lap + (a[k] / h[1] ** 2) * (p[K:K + n0, K - k:K - k + n1] + p[K:K + n0, K + k:K + k + n1])
    return lap


def _lap8(u, h):
    return lap_slices(u, h, 8)


def plain_shot_gradient(w, obs_shot, shot, v=None):
    """One shot, the slice-form step in torch ops on the GPU, recorded by autograd; loss.backward()."""
    dev = w["grid"].device
    v = torch.as_tensor(w["start"] if v is None else v, dtype=torch.float32, device=dev).requires_grad_(True)
    n, dt = w["n"], w["dt"]
    beta = torch.as_tensor(np.add.outer(w["beta"][0], w["beta"][1]), dtype=torch.float32, device=dev)
    mask = torch.zeros(n, n, device=dev)
    mask[1:-1, 1:-1] = 1.0
    kc = (v * dt) ** 2
    up = torch.zeros(n, n, device=dev)
    uc = torch.zeros(n, n, device=dev)
    si, sj = w["shots"][shot][0]
    ri = torch.as_tensor([r[0] for r in w["rec_nodes"]], device=dev)
    rj = torch.as_tensor([r[1] for r in w["rec_nodes"]], device=dev)
    sig = w["sig"][shot, :, 0].to(dev, torch.float32)
    traces = []
    for t in range(w["nsteps"]):
        num = 2 * uc - (1 - beta) * up + kc * _lap8(uc, SPACING)
        src = torch.zeros(n, n, device=dev)
        src[si, sj] = sig[t]
        num = num + kc * src
        un = num / (1 + beta) * mask
        traces.append(un[ri, rj])
        up, uc = uc, un
    d = torch.stack(traces)
    r = d - obs_shot.to(dev, torch.float32)
    J = 0.5 * (r * r).sum()
    J.backward()
    return float(J.detach()), v.grad.detach()


def timed(fn, *a, **k):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = fn(*a, **k)
    torch.cuda.synchronize()
    return out, time.perf_counter() - t0
