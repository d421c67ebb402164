"""CPU restatement of the reference's finite-difference wave stepping, in NumPy array slices.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this module; only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it,
and only as the checker.

What it restates
----------------
* ``fd_step_reference`` is ``fd_step`` of ``projects/18_simple_wave_fd/2D_wave_cupy.py:54-62``
  term for term (same slices, same operation order, NumPy scalar promotion of the Python
  float coefficients).  ``step1d_reference`` is ``step1D`` of ``step1D_wave.cu:4-17`` written
  in the same slice style.  ``reference_dt`` is ``2D_wave_cupy.py:37-39``.
* ``fd_weights`` are the central second-derivative weights of any even order in closed form;
  ``stable_dt`` is the leapfrog CFL bound built from them.  For order 2 in 2D it reduces to the
  reference's ``min(dx, dy) / c / sqrt(2)`` (``2D_wave_cupy.py:38``).
* ``laplacian`` / ``step`` / ``forward`` extend the reference update
  (``step2D_wave.cu:72-79``: ``u2 = -u0 + 2 u1 + dt^2 c^2 (Lx + Ly)`` on the interior, the
  boundary ring never written) to order 2k, a variable coefficient ``kc = v^2 dt^2``, the damped
  leapfrog sponge ``(1+b) u+ = 2u - (1-b) u- + kc (L u + s)``, point sources scaled by the same
  ``kc`` and receivers sampling ``u^{n+1}``.  Outside the grid the field is zero (the boundary
  nodes hold zero and so does everything beyond them).
* ``adjoint_gradient`` is the exact discrete transpose of ``forward`` (fp64), derived from the
  update as written: the adjoint field ``mu = kc * lambda`` obeys the forward recursion run
  backwards with the receiver residuals injected as sources of opposite sign, and
  ``dJ/dkc = -sum_n mu^{n+1} (L u^n + s^n) / kc``.

The operation order here is the reference's slice style (one second difference per axis,
divided by h^2, summed), deliberately *not* the order of the HIP kernels, so a comparison with
the kernels is a comparison of two independent writings of the same update.
"""

from __future__ import annotations

import math
from fractions import Fraction

import numpy as np


# ----------------------------------------------------------------------------- stencils
def fd_weights(order: int) -> list[Fraction]:
    """Exact central second-derivative weights ``[a_0, a_1, ..., a_K]`` for ``order = 2K``.

    ``a_k = 2 (-1)^{k+1} (K!)^2 / (k^2 (K-k)! (K+k)!)`` for ``k >= 1`` and
    ``a_0 = -2 sum_k a_k`` (the stencil annihilates constants).
    """
    if order < 2 or order % 2:
        raise ValueError(f"order must be an even number >= 2, got {order}")
    K = order // 2
    f = math.factorial
    a = [Fraction(2 * (-1) ** (k + 1) * f(K) ** 2, k * k * f(K - k) * f(K + k)) for k in range(1, K + 1)]
    return [-2 * sum(a)] + a


def stability_sum(order: int) -> Fraction:
    """``S = |a_0| + 2 sum_k |a_k|``, the largest eigenvalue magnitude of the 1D stencil."""
    a = fd_weights(order)
    return abs(a[0]) + 2 * sum(abs(x) for x in a[1:])


def stable_dt(h, vmax: float, order: int) -> float:
    """Largest stable leapfrog step: ``vmax dt <= 2 / sqrt(sum_d S / h_d^2)``."""
    S = float(stability_sum(order))
    return 2.0 / math.sqrt(sum(S / (hd * hd) for hd in h)) / vmax


def reference_dt(dx: float, dy: float, c: float, safety: float = 1.0) -> float:
    """``dt = SAFETY * min(dx, dy) / c / sqrt(2)`` (``2D_wave_cupy.py:38``)."""
    return safety * min(dx, dy) / c / math.sqrt(2)


# ------------------------------------------------------------- the reference's own step
def fd_step_reference(u0, u1, u2, dt, dx, dy, wavespeed):
    """``fd_step`` of ``2D_wave_cupy.py:54-62``, in NumPy, same slices and operation order."""
    laplacian_x = (u1[:-2, 1:-1] - 2 * u1[1:-1, 1:-1] + u1[2:, 1:-1]) / dx**2
    laplacian_y = (u1[1:-1, :-2] - 2 * u1[1:-1, 1:-1] + u1[1:-1, 2:]) / dy**2
    u2[1:-1, 1:-1] = -u0[1:-1, 1:-1] + 2 * u1[1:-1, 1:-1] + dt**2 * wavespeed**2 * (laplacian_x + laplacian_y)
    return u2


def step1d_reference(u0, u1, u2, dt, dx, wavespeed):
    """``step1D`` of ``step1D_wave.cu:4-17`` in the slice style of ``2D_wave_cupy.py:54-62``."""
    laplacian_x = (u1[:-2] - 2 * u1[1:-1] + u1[2:]) / dx**2
    u2[1:-1] = -u0[1:-1] + 2 * u1[1:-1] + dt**2 * wavespeed**2 * laplacian_x
    return u2


def run_reference(u0, u1, steps, dt, dx, dy, wavespeed):
    """The in-place two-buffer loop of ``2D_wave_cupy.py:69-71``: returns ``(u0, u1)``."""
    for _ in range(steps):
        u0 = fd_step_reference(u0, u1, u0, dt, dx, dy, wavespeed)
        u1, u0 = u0, u1
    return u0, u1


def gaussian_ic(nx, ny, lengths=(1.0, 2.0), center=(0.5, 0.5), sigma=0.02):
    """Initial pulse of ``2D_wave_cupy.py:45-50`` (float64, ``u1 = u0``)."""
    x = np.linspace(0, lengths[0], nx)
    y = np.linspace(0, lengths[1], ny)
    x, y = np.meshgrid(x, y, indexing="ij")
    return np.exp(-((x - center[0]) ** 2 + (y - center[1]) ** 2) / (2 * sigma**2))


# ------------------------------------------------------------------ general restatement
def laplacian(u: np.ndarray, order: int, h) -> np.ndarray:
    """Order-``order`` Laplacian of ``u`` on the interior nodes, zero outside the grid.

    Returned array has the shape of ``u``; boundary entries are zero.  Each axis is one
    second difference ``(a_0 u + sum_k a_k (u[+k] + u[-k])) / h_d^2`` written with slices, the
    style of ``2D_wave_cupy.py:55-56``.
    """
    dtype = u.dtype
    K = order // 2
    a = [dtype.type(float(x)) for x in fd_weights(order)]
    p = np.pad(u, K)
    inner = tuple(slice(K + 1, K + n - 1) for n in u.shape)
    out = np.zeros_like(u)
    acc = None
    for d in range(u.ndim):
        term = a[0] * p[inner]
        for k in range(1, K + 1):
            lo = list(inner)
            hi = list(inner)
            lo[d] = slice(inner[d].start - k, inner[d].stop - k)
            hi[d] = slice(inner[d].start + k, inner[d].stop + k)
            term = term + a[k] * (p[tuple(lo)] + p[tuple(hi)])
        term = term / dtype.type(h[d] ** 2)
        acc = term if acc is None else acc + term
    out[tuple(slice(1, n - 1) for n in u.shape)] = acc
    return out


def sponge_profile(n: int, width: int, beta_max: float, lo: bool = True, hi: bool = True) -> np.ndarray:
    """Quadratic damping profile ``beta(i) = beta_max ((W - e) / W)^2`` for edge distance ``e < W``.

    ``e = min(i, n-1-i)`` counts nodes from the boundary ring; only the faces flagged ``lo``
    (index 0) / ``hi`` (index n-1) get a layer.  ``beta = d dt / 2`` is the damping per step of
    the damped leapfrog, the parametrisation of ``topopt_acoustic2D.py:68-69``.
    """
    i = np.arange(n, dtype=np.float64)
    beta = np.zeros(n)
    if width > 0:
        if lo:
            e = i
            beta = np.maximum(beta, np.where(e < width, beta_max * ((width - e) / width) ** 2, 0.0))
        if hi:
            e = n - 1 - i
            beta = np.maximum(beta, np.where(e < width, beta_max * ((width - e) / width) ** 2, 0.0))
    return beta


def beta_field(profiles, dtype) -> np.ndarray:
    """``beta(x) = sum_d beta_d(x_d)``, summed slowest axis first in ``dtype``."""
    grids = np.meshgrid(*[np.asarray(p, dtype=dtype) for p in profiles], indexing="ij")
    acc = grids[0]
    for g in grids[1:]:
        acc = acc + g
    return acc


def ricker(t, f0: float, t0: float | None = None):
    """Ricker wavelet ``(1 - 2 pi^2 f0^2 tau^2) exp(-pi^2 f0^2 tau^2)``, ``tau = t - t0``, ``t0 = 1/f0``."""
    t0 = 1.0 / f0 if t0 is None else t0
    a = (math.pi * f0 * (np.asarray(t, dtype=np.float64) - t0)) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


def step(u_prev, u_cur, kc, order, h, beta=None, src_nodes=(), src_vals=()):
    """One damped leapfrog step on the interior, reference slice style; returns ``u_next``.

    ``(1 + b) u_next = 2 u_cur - (1 - b) u_prev + kc (L u_cur + s)`` with ``s`` the point
    sources (already divided by the cell volume), boundary ring left at zero.
    """
    dtype = u_cur.dtype
    lap = laplacian(u_cur, order, h)
    kc = np.broadcast_to(np.asarray(kc, dtype=dtype), u_cur.shape)
    if beta is None:
        num = 2 * u_cur - u_prev + kc * lap
    else:
        num = 2 * u_cur - (1 - beta) * u_prev + kc * lap
    for node, val in zip(src_nodes, src_vals):
        num[node] = num[node] + kc[node] * dtype.type(val)
    if beta is not None:
        num = num / (1 + beta)
    out = np.zeros_like(u_cur)
    inner = tuple(slice(1, n - 1) for n in u_cur.shape)
    out[inner] = num[inner]
    return out


def forward(kc, order, h, nsteps, beta=None, src_nodes=(), src_sig=None, rec_nodes=(),
            u_prev=None, u_cur=None, history=False):
    """Run ``nsteps`` steps from ``(u_prev, u_cur)`` (zero by default).

    Returns ``(traces[nsteps, R], u_prev, u_cur[, hist])`` with ``traces[n, r] = u^{n+1}(rec_r)``
    and, if ``history``, ``hist[n] = u^n`` for ``n = 0..nsteps-1``.  ``src_sig[n, s]`` is the
    value of source ``s`` at step ``n``; ``kc`` fixes the dtype.
    """
    kc = np.asarray(kc)
    dtype = kc.dtype
    shape = kc.shape
    u_prev = np.zeros(shape, dtype) if u_prev is None else u_prev.astype(dtype, copy=True)
    u_cur = np.zeros(shape, dtype) if u_cur is None else u_cur.astype(dtype, copy=True)
    traces = np.zeros((nsteps, len(rec_nodes)), dtype)
    hist = [] if history else None
    for n in range(nsteps):
        if history:
            hist.append(u_cur.copy())
        vals = () if src_sig is None else src_sig[n]
        u_next = step(u_prev, u_cur, kc, order, h, beta, src_nodes, vals)
        for r, node in enumerate(rec_nodes):
            traces[n, r] = u_next[node]
        u_prev, u_cur = u_cur, u_next
    if history:
        return traces, u_prev, u_cur, np.stack(hist)
    return traces, u_prev, u_cur


def misfit(traces, observed) -> float:
    """``J = 1/2 sum (d - d_obs)^2`` in float64."""
    r = np.asarray(traces, np.float64) - np.asarray(observed, np.float64)
    return 0.5 * float(np.sum(r * r))


def adjoint_gradient(kc, order, h, nsteps, observed, beta=None, src_nodes=(), src_sig=None,
                     rec_nodes=()):
    """``(J, dJ/dkc)`` by the discrete adjoint, all in float64.

    Forward: ``(1+B) u^{n+1} = 2u^n - (1-B) u^{n-1} + K (L u^n + s^n)``, ``d^n = R u^{n+1}``.
    With ``mu^m = K lambda^m`` the Lagrange conditions read
    ``(1+B) mu^m = 2 mu^{m+1} - (1-B) mu^{m+2} + K (L mu^{m+1} - R^T r^{m-1})``, i.e. ``step``
    with sources ``-r`` at the receivers, and ``dJ/dK = -sum_n mu^{n+1} (L u^n + s^n) / K``.
    """
    kc = np.asarray(kc, np.float64)
    if beta is not None:
        beta = np.asarray(beta, np.float64)
    sig = None if src_sig is None else np.asarray(src_sig, np.float64)
    traces, _, _, hist = forward(kc, order, h, nsteps, beta, src_nodes, sig, rec_nodes, history=True)
    resid = traces - np.asarray(observed, np.float64)
    J = 0.5 * float(np.sum(resid * resid))
    mu_nn = np.zeros_like(kc)  # mu^{m+2}
    mu_n = np.zeros_like(kc)  # mu^{m+1}
    G = np.zeros_like(kc)
    for m in range(nsteps, 0, -1):
        mu = step(mu_nn, mu_n, kc, order, h, beta, rec_nodes, -resid[m - 1])
        G += mu * laplacian(hist[m - 1], order, h)
        if sig is not None:
            for s, node in enumerate(src_nodes):
                G[node] += mu[node] * sig[m - 1, s]
        mu_nn, mu_n = mu_n, mu
    return J, -G / kc


# ------------------------------------------------------- storage-compressed reference steps
def fd_step_reference_f32(u0, u1, u2, dt, dx, dy, wavespeed, N, M):
    """``fd_step`` on the ``(N, M)`` nodes of fp32 arrays of any row stride (``step2D_wave.cu:59-81``
    semantics: the boundary ring and the row padding are never written)."""
    fd_step_reference(u0[:N, :M], u1[:N, :M], u2[:N, :M], dt, dx, dy, wavespeed)
    return u2


def mixed_step_reference(u0, u1, u2, dt, dx, dy, wavespeed, N, M):
    """``mixed_step`` (``step2D_wave.cu:85-124``): fp16 storage, the fp32 slice-form update, and
    round-to-nearest-even back to fp16 on the interior."""
    a0 = u0[:N, :M].astype(np.float32)
    a1 = u1[:N, :M].astype(np.float32)
    a2 = a0.copy()
    fd_step_reference(a0, a1, a2, dt, dx, dy, wavespeed)
    u2[1:N - 1, 1:M - 1] = a2[1:-1, 1:-1].astype(np.float16)
    return u2


def scaled_decode(q, s):
    """``u = s[block] * q`` with one fp32 scale per 32 consecutive entries of a row."""
    return (np.repeat(s, 32, axis=1) * q.astype(np.float32)).astype(np.float32)


def scaled_step_reference(q0, s0, q1, s1, q2, s2, dt, dx, dy, wavespeed, N, M):
    """``scaled_step`` (``step2D_wave.cu:128-183``): decode, fp32 slice-form update on the interior
    (boundary columns and padding of interior rows encode zero), re-encode each 32-node block
    with its largest magnitude m: ``q = rint(u * (32767/m))``, ``s = m / 32767``, with ``q = 0``
    where ``32767/m`` is not finite.  The reference guards ``m > 0`` (``:179``), so an all-zero
    block encodes as zero there too; it differs for ``0 < m < 32767 / FLT_MAX`` (about 9.6e-35),
    where its ``32767/m`` overflows to inf and the saturated integer conversion stores the low
    16 bits of INT_MAX (q = -1) -- here such a block stores q = 0, the value nearest to what it
    holds.  Rows 0 and N-1 are left untouched."""
    u0 = scaled_decode(q0, s0)
    u1 = scaled_decode(q1, s1)
    u2 = np.zeros_like(u1)
    fd_step_reference(u0[:N, :M], u1[:N, :M], u2[:N, :M], dt, dx, dy, wavespeed)
    rows = slice(1, N - 1)
    v = u2[rows]
    nb = v.shape[1] // 32
    blk = np.abs(v).reshape(v.shape[0], nb, 32).max(axis=2)
    inv = _block_inverse(blk)
    q = np.rint(v * np.repeat(inv, 32, axis=1)).astype(np.int16)
    q2[rows] = q
    s2[rows] = (blk / np.float32(32767.0)).astype(np.float32)
    return q2, s2


def scaled_encode(u, M):
    """Block-scaled encoding of a node array (same rule as ``scaled_step_reference``), all rows."""
    v = u.astype(np.float32)
    nb = v.shape[1] // 32
    blk = np.abs(v).reshape(v.shape[0], nb, 32).max(axis=2)
    inv = _block_inverse(blk)
    q = np.rint(v * np.repeat(inv, 32, axis=1)).astype(np.int16)
    return q, (blk / np.float32(32767.0)).astype(np.float32)


def _block_inverse(blk):
    """``32767 / m`` in fp32, zero where it is not finite (an all-zero block, or a block maximum
    below 32767 / FLT_MAX, where the quotient overflows)."""
    with np.errstate(divide="ignore", over="ignore"):
        inv = (np.float32(32767.0) / blk.astype(np.float32)).astype(np.float32)
    inv[~np.isfinite(inv)] = 0
    return inv
