"""CPU restatement of the variable mass / stiffness wave and its reconstruction adjoint.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this module; only ``tests/``
use it, as the checker of ``wfd_ac_forward`` / ``wfd_ac_backward``.

What it restates
----------------
The model is cuwave's ``AcousticWave`` as its call sites use it
(``projects/8_physics_drivers/topopt_acoustic2D.py:119-134,179-196``;
``projects/4_specialized_nns/analog_rnn_fixture.py:57-87``; ``analog_rnn_train.py:135-146``):
``m p_tt = div(s grad p) + f`` with mass ``m = 1/kappa`` and stiffness ``s = 1/rho`` per node, both
interpolated linearly in the indicator (``projects/1_benchmark_2D/README.md:141-143``), a sponge,
point sources and sensors, and gradients with respect to mass and stiffness by an adjoint that
rebuilds the forward field backwards in time from a stored boundary strip (the "boundary
reconstruction" of ``analog_rnn_fixture.py:84-87``).  cuwave's source is not available, so its
discretisation cannot be compared; this file states the one the library implements:

* ``face_q`` / ``forward_np`` / ``backward_np``: the update and the backward sweep on given
  coefficient fields ``k`` (per node) and ``cn``, ``cw`` (per face), operation for operation as
  the kernels evaluate them (same association, no fused multiply-add, fp32 or fp64), so a run
  of the library must agree bit for bit;
* ``adjoint_exact``: the transpose of the same forward map in float64 with the whole forward
  history kept (no reconstruction), the reference value of the gradient.

What pins it (tests/test_acoustic_oracle.py): energy conservation of the undamped scheme, the
closed-form cosine modes of the no-flux closure, the reduction to the constant-coefficient
5-point update, second-order convergence of the face operator on a manufactured field, the
dot-product identity of the adjoint, central finite differences, and torch autograd through a
float64 torch restatement of the forward.
"""

from __future__ import annotations

import numpy as np


def _interior(a):
    return a[1:-1, 1:-1]


def face_q(p, cn, cw):
    """``q`` on the interior nodes in the kernels' association, for K bonds per side and axis:
    ``qx = cN_1 dN_1 + cS_1 dS_1``, then ``qx = qx + (cN_m dN_m + cS_m dS_m)`` for m = 2..K, the
    same for columns, ``q = qx + qy`` (for K = 1: ``(cN dN + cS dS) + (cW dW + cE dE)``).

    ``cn`` / ``cw`` are ``(K, n0 + K - 1, n1 + K - 1)`` (a single ``(n0, n1)`` array for K = 1):
    ``cn[m-1][i, j]`` couples nodes (i-m, j) and (i, j), ``cw[m-1][i, j]`` nodes (i, j-m) and (i, j);
    outside the node box the field is zero.  Returns ``(q, (dN, dS, dW, dE))`` with lists over m."""
    cn = np.asarray(cn)
    cw = np.asarray(cw)
    if cn.ndim == 2:
        cn, cw = cn[None], cw[None]
    K = cn.shape[0]
    n0, n1 = p.shape
    P = np.zeros((n0 + 2 * K, n1 + 2 * K), p.dtype)
    P[K:K + n0, K:K + n1] = p
    C = _interior(p)
    r = slice(K + 1, K + n0 - 1)
    c = slice(K + 1, K + n1 - 1)
    dN, dS, dW, dE = [], [], [], []
    for m in range(1, K + 1):
        dN.append(P[K + 1 - m:K + n0 - 1 - m, c] - C)
        dS.append(P[K + 1 + m:K + n0 - 1 + m, c] - C)
        dW.append(P[r, K + 1 - m:K + n1 - 1 - m] - C)
        dE.append(P[r, K + 1 + m:K + n1 - 1 + m] - C)
    cN = [cn[m - 1][1:n0 - 1, 1:n1 - 1] for m in range(1, K + 1)]
    cS = [cn[m - 1][1 + m:n0 - 1 + m, 1:n1 - 1] for m in range(1, K + 1)]
    cW = [cw[m - 1][1:n0 - 1, 1:n1 - 1] for m in range(1, K + 1)]
    cE = [cw[m - 1][1:n0 - 1, 1 + m:n1 - 1 + m] for m in range(1, K + 1)]
    qx = cN[0] * dN[0] + cS[0] * dS[0]
    for m in range(1, K):
        qx = qx + (cN[m] * dN[m] + cS[m] * dS[m])
    qy = cW[0] * dW[0] + cE[0] * dE[0]
    for m in range(1, K):
        qy = qy + (cW[m] * dW[m] + cE[m] * dE[m])
    return qx + qy, (dN, dS, dW, dE)


def _bonds(a):
    a = np.asarray(a)
    return a[None] if a.ndim == 2 else a


def _damping(beta_r, beta_c, shape, dtype):
    if beta_r is None:
        return None
    br = np.asarray(beta_r, dtype=dtype)
    bc = np.asarray(beta_c, dtype=dtype)
    return br[:, None] + bc[None, :]


def _inject(num, k, pts, vals):
    """``num[x] = num[x] + k[x] * v`` for each point in list order (nodes as (i, j))."""
    for (i, j), v in zip(pts, vals):
        num[i - 1, j - 1] = num[i - 1, j - 1] + k[i, j] * v


def forward_np(k, cn, cw, nsteps, beta=(None, None), src=(), sig=None, rec=(), strip=(), dtype=np.float32,
               state=None):
    """Run ``nsteps`` steps from ``state = (p_prev, p_cur)`` (rest if None).

    Returns ``(traces [nsteps, R], store [nsteps + 1, S], (p_prev, p_cur))`` with trace row n the
    field after step n at the receivers and store row n the field before step n at the strip nodes.
    """
    dt = np.dtype(dtype).type
    k = np.asarray(k, dtype=dtype)
    cn, cw = (_bonds(a).astype(dtype) for a in (cn, cw))
    shape = k.shape
    B = _damping(beta[0], beta[1], shape, dtype)
    b = None if B is None else _interior(B)
    pp = np.zeros(shape, dtype) if state is None else np.array(state[0], dtype=dtype)
    pc = np.zeros(shape, dtype) if state is None else np.array(state[1], dtype=dtype)
    src, rec, strip = [tuple(map(tuple, np.asarray(x, np.int64).reshape(-1, 2))) for x in (src, rec, strip)]
    sig = None if sig is None else np.asarray(sig, dtype=dtype)
    traces = np.zeros((nsteps, len(rec)), dtype)
    store = np.zeros((nsteps + 1, len(strip)), dtype)
    two, one = dt(2), dt(1)
    kin = _interior(k)
    for n in range(nsteps):
        store[n] = [pc[p] for p in strip]
        q, _ = face_q(pc, cn, cw)
        C = _interior(pc)
        t = two * C - (one - b) * _interior(pp) if b is not None else two * C - _interior(pp)
        num = t + kin * q
        if src:
            _inject(num, k, src, sig[n])
        nxt = pp  # in place over the older level
        _interior(nxt)[...] = num / (one + b) if b is not None else num
        pp, pc = pc, nxt
        traces[n] = [pc[p] for p in rec]
    store[nsteps] = [pc[p] for p in strip]
    return traces, store, (pp, pc)


def undamped_mask(beta_r, beta_c, shape, dtype=np.float32):
    """Interior nodes without sponge damping."""
    m = np.zeros(shape, bool)
    if beta_r is None:
        m[1:-1, 1:-1] = True
        return m
    B = _damping(beta_r, beta_c, shape, dtype)
    m[1:-1, 1:-1] = _interior(B) == 0
    return m


def strip_nodes(beta_r, beta_c, shape, dtype=np.float32, K=1):
    """Damped interior nodes within K nodes along an axis of an undamped one, in C order, ``(S, 2)``."""
    und = undamped_mask(beta_r, beta_c, shape, dtype)
    inter = np.zeros(shape, bool)
    inter[1:-1, 1:-1] = True
    nb = np.zeros(shape, bool)
    for m in range(1, K + 1):
        nb[m:, :] |= und[:-m, :]
        nb[:-m, :] |= und[m:, :]
        nb[:, m:] |= und[:, :-m]
        nb[:, :-m] |= und[:, m:]
    return np.argwhere(inter & ~und & nb)


def backward_np(k, cn, cw, nsteps, p_last, beta=(None, None), src=(), sig=None, rec=(), adj=None, strip=(),
                store=None, dtype=np.float32):
    """The backward sweep of ``wfd_ac_backward``: ``(G [2K+1, n0+K-1, n1+K-1] float64, mu_src [N, S])``.

    ``G[0]`` the mass term on the nodes, ``G[m]`` / ``G[K+m]`` the north / west bond sums at distance
    m (bonds past the last interior row / column stored at their far end).  ``p_last = (p^{N-1},
    p^N)``; ``adj`` rows 0..N injected into mu (row n when mu^n is formed); nodes whose forward value
    the sweep does not know hold NaN, so any use of them shows.
    """
    dt = np.dtype(dtype).type
    k = np.asarray(k, dtype=dtype)
    cn, cw = (_bonds(a).astype(dtype) for a in (cn, cw))
    K = cn.shape[0]
    shape = k.shape
    n0, n1 = shape
    E0, E1 = n0 + K - 1, n1 + K - 1
    B = _damping(beta[0], beta[1], shape, dtype)
    b = None if B is None else _interior(B)
    und = undamped_mask(beta[0], beta[1], shape, dtype)
    src, rec, strip = [tuple(map(tuple, np.asarray(x, np.int64).reshape(-1, 2))) for x in (src, rec, strip)]
    sig = None if sig is None else np.asarray(sig, dtype=dtype)
    adj = None if adj is None else np.asarray(adj, dtype=dtype)
    known = und.copy()
    known[0, :] = known[-1, :] = known[:, 0] = known[:, -1] = True  # boundary ring: never written
    for p in strip:
        known[p] = True
    p_next = np.where(known, np.asarray(p_last[1], dtype), np.nan).astype(dtype)
    p_cur = np.where(known, np.asarray(p_last[0], dtype), np.nan).astype(dtype)
    mu_p = np.zeros(shape, dtype)
    mu_c = np.zeros(shape, dtype)
    part = np.zeros((2 * K + 1, E0, E1), dtype)
    G = np.zeros((2 * K + 1, E0, E1), np.float64)
    mu_src = np.zeros((nsteps, len(src)), dtype)
    two, one = dt(2), dt(1)
    kin = _interior(k)
    und_in = _interior(und)
    ni, nj = und_in.shape
    # bond at distance m to the north / west imaged when either end is an undamped interior node
    nund, wund = [], []
    for m in range(1, K + 1):
        a_ = np.zeros_like(und_in)
        a_[m:, :] = und_in[:-m, :]
        nund.append(und_in | a_)
        w_ = np.zeros_like(und_in)
        w_[:, m:] = und_in[:, :-m]
        wund.append(und_in | w_)
    rows = np.arange(1, n0 - 1)[:, None]
    cols = np.arange(1, n1 - 1)[None, :]
    imaged = 0
    for n in range(nsteps, 0, -1):
        if n <= nsteps - 1:
            mu_src[n] = [mu_c[p] for p in src]
        qm, (eN, eS, eW, eE) = face_q(mu_c, cn, cw)
        C = _interior(mu_c)
        tm = two * C - (one - b) * _interior(mu_p) if b is not None else two * C - _interior(mu_p)
        numm = tm + kin * qm
        if rec and adj is not None:
            _inject(numm, k, rec, adj[n])
        mu_n = np.zeros(shape, dtype)
        _interior(mu_n)[...] = numm / (one + b) if b is not None else numm
        if n < nsteps:
            imaged += 1
            flush = imaged % 32 == 0 or n == 1
            with np.errstate(invalid="ignore"):
                q, (dN, dS, dW, dE) = face_q(p_cur, cn, cw)
                rv = two * _interior(p_cur) + kin * q
                if src:
                    _inject(rv, k, src, sig[n])
                p_prev = np.full(shape, np.nan, dtype)
                p_prev[0, :], p_prev[-1, :], p_prev[:, 0], p_prev[:, -1] = 0, 0, 0, 0
                _interior(p_prev)[...] = np.where(und_in, rv - _interior(p_next), np.nan)
                for i, p in enumerate(strip):
                    p_prev[p] = store[n - 1, i]
                part[0, 1:n0 - 1, 1:n1 - 1] += np.where(und_in, C * q, 0).astype(dtype)
                for m in range(1, K + 1):
                    part[m, 1:n0 - 1, 1:n1 - 1] += np.where(nund[m - 1], dN[m - 1] * eN[m - 1], 0).astype(dtype)
                    part[K + m, 1:n0 - 1, 1:n1 - 1] += np.where(wund[m - 1], dW[m - 1] * eW[m - 1], 0).astype(dtype)
                    # bonds past the last interior row / column, stored at their far end
                    far_r = und_in & (rows + m >= n0 - 1)
                    ii, jj = np.nonzero(far_r)
                    part[m, ii + 1 + m, jj + 1] += (dS[m - 1] * eS[m - 1])[ii, jj].astype(dtype)
                    far_c = und_in & (cols + m >= n1 - 1)
                    ii, jj = np.nonzero(far_c)
                    part[K + m, ii + 1, jj + 1 + m] += (dE[m - 1] * eE[m - 1])[ii, jj].astype(dtype)
            if flush:
                G += part.astype(np.float64)
                part[...] = 0
            p_next, p_cur = p_cur, p_prev
        mu_p, mu_c = mu_c, mu_n
    if nsteps >= 1:
        mu_src[0] = [mu_c[p] for p in src]
    return G, mu_src


# ------------------------------------------------------------------ exact float64 transpose
def adjoint_exact(k, cn, cw, nsteps, beta=(None, None), src=(), sig=None, rec=(), dJdtr=None):
    """Gradient pieces of ``J(traces)`` in float64 with the whole forward history kept.

    ``dJdtr [N, R]`` is dJ/dtrace (trace row n = p^{n+1}).  Returns ``(dJ/dk, dJ/dcn, dJ/dcw)``:
    node array and bond arrays ``(K, n0+K-1, n1+K-1)``, imaging every interior node and every bond
    touching one.
    """
    k = np.asarray(k, np.float64)
    cn, cw = (_bonds(a).astype(np.float64) for a in (cn, cw))
    K = cn.shape[0]
    shape = k.shape
    n0, n1 = shape
    src_ = np.asarray(src, np.int64).reshape(-1, 2)
    rec_ = np.asarray(rec, np.int64).reshape(-1, 2)
    sig = None if sig is None else np.asarray(sig, np.float64)
    hist = [np.zeros(shape), np.zeros(shape)]  # p^{-1}, p^0
    B = _damping(beta[0], beta[1], shape, np.float64)
    b = None if B is None else _interior(B)
    for n in range(nsteps):
        pp, pc = hist[-2], hist[-1]
        q, _ = face_q(pc, cn, cw)
        t = 2 * _interior(pc) - ((1 - b) * _interior(pp) if b is not None else _interior(pp))
        num = t + _interior(k) * q
        for s_, (i, j) in enumerate(src_):
            num[i - 1, j - 1] += k[i, j] * sig[n, s_]
        nxt = np.zeros(shape)
        _interior(nxt)[...] = num / (1 + b) if b is not None else num
        hist.append(nxt)
    P = hist[1:]  # P[n] = p^n, n = 0..N
    mu = [np.zeros(shape), np.zeros(shape)]
    mus = {}
    for n in range(nsteps, 0, -1):
        mp, mc = mu[-2], mu[-1]
        qm, _ = face_q(mc, cn, cw)
        t = 2 * _interior(mc) - ((1 - b) * _interior(mp) if b is not None else _interior(mp))
        num = t + _interior(k) * qm
        for r, (i, j) in enumerate(rec_):
            num[i - 1, j - 1] += k[i, j] * (-dJdtr[n - 1, r])
        mn = np.zeros(shape)
        _interior(mn)[...] = num / (1 + b) if b is not None else num
        mu.append(mn)
        mus[n] = mn
    dk = np.zeros(shape)
    dcn = np.zeros(cn.shape)
    dcw = np.zeros(cw.shape)
    rows = np.arange(1, n0 - 1)[:, None]
    cols = np.arange(1, n1 - 1)[None, :]
    for n in range(nsteps):  # pairs (mu^{n+1}, p^n)
        m1, p = mus[n + 1], P[n]
        q, (dN, dS, dW, dE) = face_q(p, cn, cw)
        _, (eN, eS, eW, eE) = face_q(m1, cn, cw)
        f = np.zeros(shape)
        for s_, (i, j) in enumerate(src_):
            f[i, j] += sig[n, s_]
        with np.errstate(divide="ignore", invalid="ignore"):
            _interior(dk)[...] -= _interior(m1) * (q + _interior(f)) / _interior(k)
        for m in range(1, K + 1):
            dcn[m - 1, 1:n0 - 1, 1:n1 - 1] += dN[m - 1] * eN[m - 1]
            dcw[m - 1, 1:n0 - 1, 1:n1 - 1] += dW[m - 1] * eW[m - 1]
            ii, jj = np.nonzero(np.broadcast_to(rows + m >= n0 - 1, (n0 - 2, n1 - 2)))
            dcn[m - 1, ii + 1 + m, jj + 1] += (dS[m - 1] * eS[m - 1])[ii, jj]
            ii, jj = np.nonzero(np.broadcast_to(cols + m >= n1 - 1, (n0 - 2, n1 - 2)))
            dcw[m - 1, ii + 1, jj + 1 + m] += (dE[m - 1] * eE[m - 1])[ii, jj]
    return dk, dcn, dcw


def traces_exact(k, cn, cw, nsteps, beta=(None, None), src=(), sig=None, rec=()):
    tr, _, _ = forward_np(k, cn, cw, nsteps, beta, src, sig, rec, dtype=np.float64)
    return tr


def forward_torch(k, cn, cw, nsteps, beta=(None, None), src=(), sig=None, rec=()):
    """The forward map as differentiable torch operations (float64 tensors ``k`` and bond arrays
    ``cn``, ``cw`` of shape ``(K, n0+K-1, n1+K-1)`` or ``(n0, n1)``): traces ``[N, R]``."""
    import torch
    import torch.nn.functional as F

    if cn.dim() == 2:
        cn, cw = cn[None], cw[None]
    K = cn.shape[0]
    shape = k.shape
    n0, n1 = shape
    src_ = torch.as_tensor(np.asarray(src, np.int64).reshape(-1, 2))
    rec_ = torch.as_tensor(np.asarray(rec, np.int64).reshape(-1, 2))
    sig = None if sig is None else torch.as_tensor(sig, dtype=torch.float64)
    b = None
    if beta[0] is not None:
        br = torch.as_tensor(np.asarray(beta[0], np.float64))
        bc = torch.as_tensor(np.asarray(beta[1], np.float64))
        b = (br[:, None] + bc[None, :])[1:-1, 1:-1]
    pp = torch.zeros(shape, dtype=torch.float64)
    pc = torch.zeros(shape, dtype=torch.float64)
    kin = k[1:-1, 1:-1]
    out = []
    for n in range(nsteps):
        P = F.pad(pc, (K, K, K, K))
        C = pc[1:-1, 1:-1]
        q = 0
        for m in range(1, K + 1):
            q = q + cn[m - 1][1:n0 - 1, 1:n1 - 1] * (P[K + 1 - m:K + n0 - 1 - m, K + 1:K + n1 - 1] - C) \
                + cn[m - 1][1 + m:n0 - 1 + m, 1:n1 - 1] * (P[K + 1 + m:K + n0 - 1 + m, K + 1:K + n1 - 1] - C) \
                + cw[m - 1][1:n0 - 1, 1:n1 - 1] * (P[K + 1:K + n0 - 1, K + 1 - m:K + n1 - 1 - m] - C) \
                + cw[m - 1][1:n0 - 1, 1 + m:n1 - 1 + m] * (P[K + 1:K + n0 - 1, K + 1 + m:K + n1 - 1 + m] - C)
        t = 2 * C - ((1 - b) * pp[1:-1, 1:-1] if b is not None else pp[1:-1, 1:-1])
        f = torch.zeros_like(C)
        if len(src_):
            f = f.index_put((src_[:, 0] - 1, src_[:, 1] - 1), sig[n], accumulate=True)
        num = t + kin * (q + f)
        nxt = F.pad(num / (1 + b) if b is not None else num, (1, 1, 1, 1))
        pp, pc = pc, nxt
        out.append(pc[rec_[:, 0], rec_[:, 1]])
    return torch.stack(out)
