"""Torch-CPU restatement of the forward run, differentiated by torch autograd.

TEST INFRASTRUCTURE ONLY (tests/ use it as the independent gradient check of the hand-written
adjoint).  The step is ``slice_ref.step`` written with torch ops (reference slice style of
``2D_wave_cupy.py:54-62``), so ``torch.autograd`` gives dJ/dv through the whole time loop.
"""

from __future__ import annotations

import torch
import torch.nn.functional as F

from .slice_ref import fd_weights


def laplacian(u: torch.Tensor, order: int, h) -> torch.Tensor:
    K = order // 2
    a = [float(x) for x in fd_weights(order)]
    nd = u.dim()
    p = F.pad(u, [K] * (2 * nd))
    inner = tuple(slice(K + 1, K + n - 1) for n in u.shape)
    acc = None
    for d in range(nd):
        term = a[0] * p[inner]
        for k in range(1, K + 1):
            lo = list(inner)
            hi = list(inner)
            lo[d] = slice(inner[d].start - k, inner[d].stop - k)
            hi[d] = slice(inner[d].start + k, inner[d].stop + k)
            term = term + a[k] * (p[tuple(lo)] + p[tuple(hi)])
        term = term / h[d] ** 2
        acc = term if acc is None else acc + term
    return F.pad(acc, [1] * (2 * nd))


def forward_traces(v, dt, order, h, nsteps, beta=None, src_nodes=(), src_sig=None, rec_nodes=()):
    """Traces ``d[n, r] = u^{n+1}(rec_r)`` for velocity ``v`` (differentiable)."""
    kc = (v * dt) ** 2
    u_prev = torch.zeros_like(v)
    u_cur = torch.zeros_like(v)
    interior = torch.zeros_like(v)
    interior[tuple(slice(1, n - 1) for n in v.shape)] = 1.0
    rec_idx = tuple(torch.as_tensor([r[d] for r in rec_nodes]) for d in range(v.dim()))
    out = []
    for n in range(nsteps):
        lap = laplacian(u_cur, order, h)
        if beta is None:
            num = 2 * u_cur - u_prev + kc * lap
        else:
            num = 2 * u_cur - (1 - beta) * u_prev + kc * lap
        if src_sig is not None and len(src_nodes):
            inj = torch.zeros_like(v)
            for s, node in enumerate(src_nodes):
                inj = inj.index_put((torch.tensor([node[0]]), torch.tensor([node[1]]))
                                    if v.dim() == 2 else tuple(torch.tensor([c]) for c in node),
                                    torch.as_tensor([float(src_sig[n][s])], dtype=v.dtype),
                                    accumulate=True)
            num = num + kc * inj
        if beta is not None:
            num = num / (1 + beta)
        u_next = num * interior
        out.append(u_next[rec_idx])
        u_prev, u_cur = u_cur, u_next
    return torch.stack(out)


def misfit_grad_v(v, dt, order, h, nsteps, observed, **kw):
    """``(J, dJ/dv)`` through torch autograd, ``J = 1/2 sum (d - d_obs)^2``."""
    v = v.detach().clone().requires_grad_(True)
    d = forward_traces(v, dt, order, h, nsteps, **kw)
    J = 0.5 * ((d - observed) ** 2).sum()
    J.backward()
    return float(J), v.grad.detach()
