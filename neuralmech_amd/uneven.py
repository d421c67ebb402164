"""The uneven batch workload: 64 shots drawn from ``default_rng(4)`` -- square grids of side
n ~ U{512..1536} nodes, step counts ~ U{1500..5000}, order 4 or 8 -- each with the FWI model pair of
that size (layered true model with a seeded +-10 % perturbation, start model = true smoothed with
sigma 40 nodes), dx = 10 m, 40-node sponge, one Ricker 10 Hz source and 201 receivers on row 45.
The observed records come from the true models; every shot's gradient is taken at its start model.

``batched`` runs all 64 gradients in one persistent launch (:class:`~neuralmech_amd.batch.ShotBatch`);
``sequential`` runs the same shots one by one through the single-shot gradient plan (one launch
per step); ``plain_shot_gradient`` is the plain per-shot forward-then-autograd form (the slice-form
step in torch ops, recorded by autograd, then ``loss.backward()``), the per-shot loop of
``projects/18_fwi/2D_fwi.py:39-41`` written in torch.
"""

from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from . import media, solver
from .batch import Shot, ShotBatch
from .fwi import lap_slices

SPACING = (10.0, 10.0)


def draw(nshots=64, seed=4):
    """Per shot ``(n, nsteps, order, source column)`` from ``default_rng(seed)``."""
    rng = np.random.default_rng(seed)
    n = rng.integers(512, 1537, nshots)
    steps = rng.integers(1500, 5001, nshots)
    order = rng.choice([4, 8], nshots)
    col = [int(rng.integers(64, int(k) - 64)) for k in n]
    return [(int(a), int(b), int(c), d) for a, b, c, d in zip(n, steps, order, col)]


def setup(nshots=64, seed=4, dtype=torch.float32, step_scale=1.0, threads=8):
    """The shots with observed records (forward runs of the true models on the GPU)."""
    spec = draw(nshots, seed)
    with ThreadPoolExecutor(threads) as ex:
        models = list(ex.map(lambda k: media.layered_fwi_models(spec[k][0], seed=1000 * seed + k), range(nshots)))
    shots = []
    for k, ((n, N, order, col), (true, start)) in enumerate(zip(spec, models)):
        N = max(2, int(round(N * step_scale)))
        dt = 0.9 * solver.stable_dt(SPACING, float(max(true.max(), start.max())), order)
        src = [(45, col)]
        rec = [(45, int(c)) for c in np.round(np.linspace(41, n - 42, 201)).astype(int)]
        sig = media.ricker(np.arange(N) * dt, 10.0) / (SPACING[0] * SPACING[1])
        beta = [media.sponge_profile(n, 40, 0.05) for _ in range(2)]
        g = solver.WaveGrid((n, n), SPACING, dt, order, dtype)
        med = solver.Medium(g, velocity=torch.as_tensor(true), beta_profiles=beta)
        tr, _ = solver.run(g, med, N, src=solver.PointSet(g, src), sig=torch.as_tensor(sig)[None, :, None],
                           rec=solver.PointSet(g, rec))
        shots.append(Shot((n, n), SPACING, dt, order, start, src, sig[:, None], rec, tr[0].cpu(), N, beta))
        del med, g, tr
    return shots


def cell_steps(shots) -> int:
    return sum((s.shape[0] - 2) * (s.shape[1] - 2) * s.nsteps for s in shots)


def batched(shots, dtype=torch.float32, tune_steps=10, **kw):
    """Build (and tune) the batch; returns the ShotBatch ready to run."""
    b = ShotBatch(shots, dtype, **kw)
    if tune_steps:
        b.tune(tune_steps)
    return b


def sequential(shots, dtype=torch.float32):
    """Every shot through the single-shot gradient plan, one after another; returns the gradients
    and the elapsed seconds (device-synchronised)."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = []
    for sh in shots:
        g = solver.WaveGrid(sh.shape, sh.spacing, sh.dt, sh.order, dtype)
        med = solver.Medium(g, velocity=torch.as_tensor(sh.velocity), beta_profiles=sh.beta_profiles)
        src = solver.PointSet(g, sh.src_nodes)
        rec = solver.PointSet(g, sh.rec_nodes)
        plan = solver.GradientPlan(g, sh.nsteps, seg=sh.seg)
        sig = torch.as_tensor(sh.signal)[None].to("cuda", dtype).contiguous()
        tr = plan.forward(med, src, sig, rec)
        resid = tr - torch.as_tensor(sh.observed)[None].to("cuda", dtype)
        grad_kc = plan.backward(med, src, rec, resid)
        v = torch.as_tensor(sh.velocity, dtype=torch.float64, device="cuda")
        out.append(grad_kc * 2.0 * v * g.dt ** 2)
        del plan
    torch.cuda.synchronize()
    return out, time.perf_counter() - t0


def plain_shot_gradient(sh: Shot, dtype=torch.float32):
    """``(J, dJ/dv)`` of one shot, the plain way: the shot's slice-form step (its order, sponge,
    source divided by the cell volume in the signal) in torch ops on the GPU, recorded by autograd
    over every step, then ``loss.backward()`` with ``J = 1/2 sum (d - d_obs)^2``.  The damping
    factors are formed once, outside the step loop, so the tape holds about one field per step
    (the largest shot, 1536^2 x 5000 steps, needs ~50 GB)."""
    dev = torch.device("cuda")
    n0, n1 = sh.shape
    v = torch.as_tensor(sh.velocity, dtype=dtype, device=dev).requires_grad_(True)
    if sh.beta_profiles is not None:
        beta = torch.as_tensor(np.add.outer(sh.beta_profiles[0], sh.beta_profiles[1]), dtype=dtype, device=dev)
    else:
        beta = torch.zeros(n0, n1, dtype=dtype, device=dev)
    keep, inv = 1 - beta, 1 / (1 + beta)
    mask = torch.zeros(n0, n1, dtype=dtype, device=dev)
    mask[1:-1, 1:-1] = 1.0
    kc = (v * sh.dt) ** 2
    up = torch.zeros(n0, n1, dtype=dtype, device=dev)
    uc = torch.zeros(n0, n1, dtype=dtype, device=dev)
    (si, sj), = [tuple(int(c) for c in p) for p in sh.src_nodes]
    ri = torch.as_tensor([int(r[0]) for r in sh.rec_nodes], device=dev)
    rj = torch.as_tensor([int(r[1]) for r in sh.rec_nodes], device=dev)
    sig = torch.as_tensor(np.asarray(sh.signal).reshape(sh.nsteps, -1)[:, 0], dtype=dtype, device=dev)
    hit = torch.zeros(n0, n1, dtype=dtype, device=dev)
    hit[si, sj] = 1.0
    traces = []
    for t in range(sh.nsteps):
        num = 2 * uc - keep * up + kc * (lap_slices(uc, sh.spacing, sh.order) + hit * sig[t])
        un = num * inv * mask
        traces.append(un[ri, rj])
        up, uc = uc, un
    d = torch.stack(traces)
    r = d - torch.as_tensor(sh.observed, dtype=dtype, device=dev)
    J = 0.5 * (r * r).sum()
    J.backward()
    return float(J.detach()), v.grad.detach()


def plain_sample(shots, k: int = 3):
    """Indices of ``k`` shots spread over the batch's range of cell-steps (smallest, ..., largest)."""
    cs = [(s.shape[0] - 2) * (s.shape[1] - 2) * s.nsteps for s in shots]
    order = sorted(range(len(shots)), key=lambda i: cs[i])
    if k <= 1:
        return [order[len(order) // 2]]
    return sorted({order[round(j * (len(order) - 1) / (k - 1))] for j in range(k)})
