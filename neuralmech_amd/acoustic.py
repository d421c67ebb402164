"""Variable mass / stiffness acoustic wave over ``wfd_ac_forward`` / ``wfd_ac_backward``.

Replaces cuwave's ``AcousticWave`` and its sensitivities as the drivers call them:
``AcousticWave(Nx, dx, N, dt, threads, precision, space_order, rho1, rho2, kappa1, kappa2)`` and
``parametrization_jacobian`` (``projects/8_physics_drivers/topopt_acoustic2D.py:119-131``,
``projects/4_specialized_nns/analog_rnn_train.py:136``), ``response`` / ``response_gradient`` /
``energy`` (``topopt_acoustic2D.py:167,179-181,194-196``), ``reconstruction_sensitivity`` and
``reconstruction_nodes`` (``analog_rnn_train.py:142-146``, ``analog_rnn_fixture.py:84-87``).

Model and discretisation (cuwave's source is not available, so these are this build's):

* ``m p_tt = div(s grad p) + f`` with mass ``m = (1 - g)/kappa1 + g/kappa2`` and stiffness
  ``s = (1 - g)/rho1 + g/rho2`` per node for an indicator ``g`` in [0, 1]: 1/rho and 1/kappa
  interpolate, so a grey value mixes harmonically (``projects/1_benchmark_2D/README.md:141-143``);
  the wave speed is ``sqrt(s / m)``.
* Finite volumes on the nodes.  A face between two nodes carries ``s_f / h^2`` with ``s_f`` the
  harmonic mean of the two nodes' stiffness (the reciprocal of the mean density), which keeps
  the explicit step stable across the high-contrast interfaces of a binary design.
* Closure ``"neumann"`` (no flux; the edges reflect, ``projects/1_benchmark_2D/forward/wave.py:44``):
  faces to the boundary ring carry nothing, the first / last interior node of an axis owns half a
  control volume and faces along a wall half an area (bilinear finite volumes with lumped mass), so
  the walls sit on those nodes, ``dx = L / (n - 3)`` as at the call sites.
  Closure ``"dirichlet"``: the ring holds zero, every face to it is active (the scalar model's).
* Leapfrog with the sponge of the scalar model, ``(1+b) p+ = 2p - (1-b) p- + k (C p + f)``,
  ``k = w dt^2 / m`` (``w`` the control-volume factor), sources divided by the cell volume.
* Gradients by the hand-written adjoint with boundary reconstruction: the forward field is
  rebuilt backwards in time from (p^{N-1}, p^N) where nothing damps it, and a strip of damped
  nodes next to that region is recorded every step, ``(N+1) x strip`` values per shot.  Gradients
  are therefore available on the interior nodes free of sponge damping and are zero elsewhere.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np
import torch
import torch.nn.functional as F

from . import _lib, solver
from ._lib import check
from .solver import PointSet, WaveGrid, _ptr, _stream_ptr

CLOSURES = ("neumann", "dirichlet")
_PRECISIONS = {"float32": torch.float32, "float64": torch.float64}


# ----------------------------------------------------------------------- coefficients
def volume_weights(shape, closure: str) -> np.ndarray:
    """Control-volume factor ``w`` per node: 2 per axis along which a node is the first or last
    interior node under the no-flux closure, 1 otherwise."""
    n0, n1 = shape
    w = np.ones(shape)
    if closure == "neumann":
        for i in (1, n0 - 2):
            w[i, :] *= 2.0
        for j in (1, n1 - 2):
            w[:, j] *= 2.0
    return w


def central_weights(order: int) -> list:
    """``a_1 .. a_K`` of the central second difference of ``order`` (closed form,
    ``a_k = 2 (-1)^(k+1) (K!)^2 / (k^2 (K-k)! (K+k)!)``)."""
    K = order // 2
    return [2.0 * (-1) ** (k + 1) * math.factorial(K) ** 2 / (k * k * math.factorial(K - k) * math.factorial(K + k))
            for k in range(1, K + 1)]


def _harm(a, b):
    return 2.0 * a * b / (a + b)


POWER_MEAN_Q = 8  # exponent -q of the segment mean of the stiffness on bonds of orders >= 4


def _segment_mean(sp, lo, hi, q: int):
    """Power mean ``((1/n) sum s^-q)^(-1/q)`` of ``sp`` over rows ``lo[b] .. hi[b]`` (inclusive, all
    columns): rows gathered per bond, scaled by the largest value first so that ``s^-q`` stays in
    range (the mean is homogeneous of degree 1, so the scale passes through)."""
    ref = sp.detach().max()
    n = (hi - lo + 1).to(sp.dtype)[:, None]
    acc = torch.zeros((len(lo), sp.shape[1]), dtype=sp.dtype, device=sp.device)
    for t in range(int((hi - lo).max()) + 1):
        rows = torch.minimum(lo + t, hi)
        valid = (lo + t <= hi).to(sp.dtype)[:, None]
        acc = acc + valid * (sp[rows] / ref) ** (-q)
    return ref * (acc / n) ** (-1.0 / q)


def _axis_bonds(s, h: float, order: int, closure: str):
    """Bond coefficients along axis 0 of the node array ``s``: ``K`` arrays of shape
    ``(n0 + K - 1, n1)``, entry ``[m-1][i, j]`` the bond between (i-m, j) and (i, j).

    Order 2: the harmonic mean of the two ends.  Orders 4..8: the power mean with exponent -8 of
    every node on the segment between the ends (a reflected bond's segment runs through the wall);
    the negative bonds of these stencils (a_2, a_4 < 0) then never outweigh the shorter bonds they
    span (their segment mean is at most (3/2)^(1/8) of any sub-segment's), which keeps the operator
    negative semi-definite at any contrast; with constant stiffness the mean is that constant.
    Zero ring: every bond with an interior end, the stiffness beyond the ring continued from it.
    No flux: bonds between interior nodes, plus the bonds the even reflection of the field about the
    first / last interior node adds near the walls (a node's bond to a mirror image of another node
    becomes a bond to that node; bonds to its own image drop out), so that the operator is the
    unreflected stencil applied to the even extension.
    """
    K = order // 2
    a = central_weights(order)
    n0, n1 = s.shape
    E0 = n0 + K - 1
    cols = torch.arange(1, n1 - 1, device=s.device)[None, :]
    dev = s.device

    def coef(rows_n, rows_s, lo, hi, sp, off, w):
        # bonds between rows_n and rows_s (node rows of sp shifted by off), segment rows lo..hi
        if K == 1:
            v = _harm(sp[rows_n + off][:, 1:n1 - 1], sp[rows_s + off][:, 1:n1 - 1])
        else:
            v = _segment_mean(sp[:, 1:n1 - 1], lo + off, hi + off, POWER_MEAN_Q)
        return w[:, None] / (h * h) * v

    out = []
    if closure == "dirichlet":
        sp = torch.cat([s[:1].expand(K, n1), s, s[-1:].expand(K, n1)], 0)  # row i at i + K
        for m in range(1, K + 1):
            i = torch.arange(1, n0 - 1 + m, device=dev)
            w = torch.full((len(i),), a[m - 1], dtype=s.dtype, device=dev)
            v = coef(i - m, i, i - m, i, sp, K, w)
            f = torch.zeros((E0, n1), dtype=s.dtype, device=dev)
            out.append(f.index_put((i[:, None], cols), v))
        return out
    L = n0 - 2  # last interior row; the first is 1
    entries = {m: [] for m in range(1, K + 1)}  # distance -> [(southern row, northern row, weight, lo, hi)]
    for m in range(1, K + 1):
        for i in range(1 + m, L + 1):
            entries[m].append((i, i - m, a[m - 1], i - m, i))
    for i in range(2, min(K, L) + 1):  # near the first interior row: images 2 - (i - m)
        for m in range(i, K + 1):
            im = 2 - i + m
            if im > i and im <= L:
                entries[im - i].append((im, i, a[m - 1], 1, im))
    for i in range(max(L - K + 1, 1), L):  # near the last interior row: images 2L - (i + m)
        for m in range(L + 1 - i, K + 1):
            im = 2 * L - i - m
            if 1 <= im < i:
                entries[i - im].append((i, im, a[m - 1], im, L))
    for d in range(1, K + 1):
        f = torch.zeros((E0, n1), dtype=s.dtype, device=dev)
        if entries[d]:
            col = lambda k, dt=torch.long: torch.as_tensor([e[k] for e in entries[d]], dtype=dt, device=dev)  # noqa: E731
            rs, rn, lo, hi = col(0), col(1), col(3), col(4)
            w = col(2, s.dtype)
            v = coef(rn, rs, lo, hi, s, 0, w)
            f = f.index_put((rs[:, None], cols), v, accumulate=True)
        out.append(f)
    return out


def face_operator(mass: torch.Tensor, stiff: torch.Tensor, spacing, dt: float, closure: str, order: int = 2):
    """``(k, cn, cw)``: ``k`` on the nodes and the bond coefficients ``(K, n0+K-1, n1+K-1)``
    (float64, differentiable in ``mass`` and ``stiff``); K = order / 2.

    The bond between nodes x and y at distance m along an axis is ``a_m hm(s_x, s_y) / h^2`` times
    its relative area (the harmonic mean ``hm`` is the reciprocal of the mean density); with
    constant stiffness the operator is the central stencil of ``order``.  Under the no-flux closure
    a bond along a wall (both ends on the first / last interior row or column) has half the area and
    ``k = w dt^2 / m`` with ``w`` = 2 per axis on those rows / columns (half control volumes)."""
    if closure not in CLOSURES:
        raise ValueError(f"closure must be one of {CLOSURES}")
    if order not in (2, 4, 6, 8):
        raise ValueError("order must be 2, 4, 6 or 8")
    K = order // 2
    shape = tuple(mass.shape)
    n0, n1 = shape
    dev = mass.device
    inter = torch.zeros(shape, dtype=torch.bool, device=dev)
    inter[1:-1, 1:-1] = True
    w = torch.as_tensor(volume_weights(shape, closure), device=dev)
    k = torch.where(inter, dt * dt * w / mass, torch.zeros((), dtype=mass.dtype, device=dev))
    h0, h1 = (float(h) for h in spacing)
    bn = _axis_bonds(stiff, h0, order, closure)        # (n0+K-1, n1) each
    bw = _axis_bonds(stiff.t(), h1, order, closure)    # (n1+K-1, n0) each, transposed
    if closure == "neumann":
        an = torch.ones(n1, dtype=stiff.dtype, device=dev)
        an[1] = an[n1 - 2] = 0.5
        aw = torch.ones(n0, dtype=stiff.dtype, device=dev)
        aw[1] = aw[n0 - 2] = 0.5
        bn = [f * an[None, :] for f in bn]
        bw = [f * aw[None, :] for f in bw]
    cn = torch.stack([F.pad(f, (0, K - 1, 0, 0)) for f in bn])
    cw = torch.stack([F.pad(f.t(), (0, 0, 0, K - 1)) for f in bw])
    return k, cn, cw


def _profiles(grid: WaveGrid, beta_profiles):
    """Sponge profiles rounded to the grid dtype, as host float64 arrays (or None)."""
    if beta_profiles is None:
        return None
    if len(beta_profiles) != 2:
        raise ValueError("one damping profile per axis")
    npdt = np.float32 if grid.dtype == torch.float32 else np.float64
    out = []
    for n, p in zip(grid.shape, beta_profiles):
        p = np.asarray(torch.as_tensor(p, dtype=torch.float64).cpu(), np.float64).reshape(-1)
        if p.size != n:
            raise ValueError("damping profile length must equal the axis node count")
        if not np.all(np.isfinite(p)) or np.any(p < 0) or np.any(p >= 1):
            raise ValueError("damping profiles must be finite and in [0, 1)")
        out.append(p.astype(npdt).astype(np.float64))
    return out


def ext_view(grid: WaveGrid, field: torch.Tensor, K: int) -> torch.Tensor:
    """``(b, n0+K-1, n1+K-1)`` view of layout tensors ``(b, elems)``: the nodes plus the K-1 rows and
    columns past the last node that hold bonds reaching beyond the last interior row / column."""
    b = field.shape[0]
    n0, n1 = grid.shape
    o = grid.offset
    rows = n0 + K - 1
    return field[:, o:o + rows * grid.ld].view(b, rows, grid.ld)[:, :, : n1 + K - 1]


class AcMedium:
    """Coefficient fields of the variable mass / stiffness model on a 2D grid (order 2..8).

    ``mass`` / ``stiff`` are node arrays (or numbers) shared by all shots; ``closure`` is
    ``"neumann"`` or ``"dirichlet"``; ``beta_profiles`` one damping profile per axis.
    """

    def __init__(self, grid: WaveGrid, mass, stiff, closure: str = "neumann", beta_profiles=None):
        if grid.ndim != 2:
            raise ValueError("the variable mass / stiffness model runs on 2D grids")
        if closure not in CLOSURES:
            raise ValueError(f"closure must be one of {CLOSURES}")
        K = grid.order // 2
        if closure == "neumann" and min(grid.shape) < 2 * K + 3:
            raise ValueError(f"the no-flux closure of order {grid.order} needs at least {2 * K + 3} nodes per axis")
        self.grid, self.closure, self.K = grid, closure, K
        dev = grid.device
        self.mass = torch.as_tensor(mass, dtype=torch.float64).to(dev).expand(grid.shape).clone()
        self.stiff = torch.as_tensor(stiff, dtype=torch.float64).to(dev).expand(grid.shape).clone()
        for name, a in (("mass", self.mass), ("stiffness", self.stiff)):
            if not bool(torch.isfinite(a).all()) or not bool((a > 0).all()):
                raise ValueError(f"{name} must be finite and positive on every node")
        k, cn, cw = face_operator(self.mass, self.stiff, grid.spacing, grid.dt, closure, grid.order)
        self.k = grid.from_nodes(k, nshots=1, dtype=grid.dtype)
        self.cn = grid.zeros(nshots=K)
        self.cw = grid.zeros(nshots=K)
        ext_view(grid, self.cn, K).copy_(cn)
        ext_view(grid, self.cw, K).copy_(cw)
        prof = _profiles(grid, beta_profiles)
        self.profiles = prof
        self.beta = None
        if prof is not None:
            self.beta = torch.as_tensor(np.concatenate(prof), dtype=grid.dtype).to(dev)
        und = np.zeros(grid.shape, bool)
        if prof is None:
            und[1:-1, 1:-1] = True
        else:
            und[1:-1, 1:-1] = (prof[0][:, None] + prof[1][None, :])[1:-1, 1:-1] == 0
        self.undamped = und
        self.c = _lib.AcMedium()
        self.c.k, self.c.cn, self.c.cw = self.k.data_ptr(), self.cn.data_ptr(), self.cw.data_ptr()
        self.c.beta = None if self.beta is None else self.beta.data_ptr()

    def bonds(self):
        """``(k, cn, cw)`` as the kernels use them (grid dtype): node array and ``(K, n0+K-1, n1+K-1)``."""
        g = self.grid
        return g.nodes(self.k)[0], ext_view(g, self.cn, self.K), ext_view(g, self.cw, self.K)

    def strip_nodes(self) -> np.ndarray:
        """Damped interior nodes within K nodes along an axis of an undamped one, ``(S, 2)``, C order."""
        return _strip(self.undamped, self.K)

    def check_step(self) -> float:
        """Gershgorin bound of ``dt^2 lambda_max`` of the update (stable below 4)."""
        k, cn, cw = (a.double() for a in self.bonds())
        n0, n1 = self.grid.shape
        row = torch.zeros_like(k)
        for m in range(1, self.K + 1):
            row += cn[m - 1][:n0, :n1] + cn[m - 1][m:m + n0, :n1]
            row += cw[m - 1][:n0, :n1] + cw[m - 1][:n0, m:m + n1]
        return float((2.0 * k * row).max())


def _strip(und: np.ndarray, K: int) -> np.ndarray:
    nb = np.zeros_like(und)
    for m in range(1, K + 1):
        nb[m:, :] |= und[:-m, :]
        nb[:-m, :] |= und[m:, :]
        nb[:, m:] |= und[:, :-m]
        nb[:, :-m] |= und[:, m:]
    inter = np.zeros_like(und)
    inter[1:-1, 1:-1] = True
    return np.argwhere(inter & ~und & nb)


# ------------------------------------------------------------------------- runs
def set_rows(rows: int = 0) -> None:
    """Rows per wave of the variable mass / stiffness kernels for every later launch in this
    process: 0 lets each launch choose, 1 / 2 / 4 / 8 force that count (results are the same
    bits whatever it is)."""
    check(_lib.load().wfd_ac_set_rows(int(rows)))


def forward(grid: WaveGrid, med: AcMedium, nsteps: int, src: Optional[PointSet] = None, sig=None,
            rec: Optional[PointSet] = None, strip: Optional[PointSet] = None, state=None, step0=None):
    """Advance ``nsteps`` steps: ``(traces, store, state)``.

    ``traces[s, n, r]`` = p^{step0+n+1} at receiver r; ``store[s, n, i]`` = p^{step0+n} at strip node
    i for n = 0..nsteps.  ``state`` is a ``solver.WaveState`` (``prev = p^{n-1}``, ``cur = p^n``,
    ``step = n``): rest at ``step0`` (default 0) when omitted, otherwise continued from
    ``state.step`` and advanced in place, levels and step, so that chained calls equal one call.
    """
    lib = _lib.load()
    state = solver._as_state(grid, state, step0)
    step0 = state.step
    ua, ub = state.prev, state.cur
    total = step0 + nsteps
    s_sig = solver._sig_tensor(grid, src, sig, total)
    traces = None
    if rec is not None and rec.count:
        traces = torch.zeros(grid.nshots, total, rec.count, dtype=grid.dtype, device=grid.device)
    store = None
    if strip is not None and strip.count:
        store = torch.zeros(grid.nshots, total + 1, strip.count, dtype=grid.dtype, device=grid.device)
    sp = src.struct(s_sig) if src is not None else None
    rp = rec.struct(traces) if rec is not None else None
    tp = strip.struct(store) if strip is not None else None
    check(lib.wfd_ac_forward(ctypes.byref(grid.c), _ptr(ua), _ptr(ub), ctypes.byref(med.c),
                             ctypes.byref(sp) if sp else None, ctypes.byref(rp) if rp else None,
                             ctypes.byref(tp) if tp else None, int(step0), int(nsteps), _stream_ptr(grid.device)))
    if nsteps % 2:  # the newest level is now in the buffer that held p^{n-1}
        state.prev, state.cur = state.cur, state.prev
    state.step += int(nsteps)
    tr = None if traces is None else traces[:, step0:]
    st = None if store is None else store[:, step0:]
    return tr, st, state


def backward(grid: WaveGrid, med: AcMedium, nsteps: int, state, src: Optional[PointSet], sig, rec: PointSet,
             dJdtr, strip: Optional[PointSet] = None, store=None, noalloc: bool = False):
    """Adjoint sweep with reconstruction for ``J(traces)`` with ``dJ/dtraces = dJdtr``
    ``(nshots, nsteps, R)``; ``state = (p^{N-1}, p^N)`` from :func:`forward` (consumed).

    Returns ``(G, mu_src)``: ``G`` float64 ``(2K+1, nshots, n0+K-1, n1+K-1)`` (mass term on the nodes,
    then the north and the west bond sums per distance; see ``wfd_ac_backward``) and ``mu_src``
    ``(nshots, nsteps, S)`` or None.  ``noalloc`` runs ``wfd_ac_backward_noalloc``: the library's
    work fields come from its per-device scratch array (at most ``WFD_AC_SCRATCH_BYTES``), the same
    bits; calls on one device must not overlap.
    """
    lib = _lib.load()
    g = grid
    dJ = torch.as_tensor(dJdtr, device=g.device)
    if dJ.shape != (g.nshots, nsteps, rec.count):
        raise ValueError(f"dJ/dtraces must be (nshots, nsteps, receivers) = {(g.nshots, nsteps, rec.count)}")
    adj = torch.zeros(g.nshots, nsteps + 1, rec.count, dtype=g.dtype, device=g.device)
    adj[:, 1:] = (-dJ).to(g.dtype)
    s_sig = solver._sig_tensor(g, src, sig, nsteps)
    sp = src.struct(s_sig) if src is not None else None
    mu_src = None
    if src is not None and src.count:
        mu_src = torch.zeros(g.nshots, nsteps, src.count, dtype=g.dtype, device=g.device)
    mp = src.struct(mu_src) if mu_src is not None else None
    ap = rec.struct(adj)
    if strip is not None and strip.count and store is None:
        raise ValueError("the strip needs the values the forward recorded")
    tp = strip.struct(store.contiguous()) if strip is not None and strip.count else None
    nf = 2 * med.K + 1
    G = torch.zeros(nf, g.nshots, g.elems, dtype=torch.float64, device=g.device)
    ua, ub = state
    pts = (ctypes.byref(sp) if sp else None, ctypes.byref(ap) if ap else None,
           ctypes.byref(tp) if tp else None, ctypes.byref(mp) if mp else None)
    if noalloc:  # work fields in the library's per-device scratch array
        check(lib.wfd_ac_backward_noalloc(ctypes.byref(g.c), _ptr(ua), _ptr(ub), ctypes.byref(med.c), *pts,
                                          int(nsteps), _ptr(G), _stream_ptr(g.device)))
    else:
        mu_a, mu_b = g.zeros(), g.zeros()
        work = torch.zeros(nf * g.nshots * g.elems, dtype=g.dtype, device=g.device)
        check(lib.wfd_ac_backward(ctypes.byref(g.c), _ptr(ua), _ptr(ub), _ptr(mu_a), _ptr(mu_b),
                                  ctypes.byref(med.c), *pts, int(nsteps), _ptr(work), _ptr(G),
                                  _stream_ptr(g.device)))
    return torch.stack([ext_view(g, G[a], med.K) for a in range(nf)]), mu_src


def material_gradient(med: AcMedium, G: torch.Tensor, mu_src, sig, src: Optional[PointSet],
                      sum_shots: bool = False):
    """``(dJ/dmass, dJ/dstiff)`` per shot ``(nshots, n0, n1)`` float64 from the sweep's sums ``G``
    ``(2K+1, nshots, n0+K-1, n1+K-1)``: ``dJ/dk = -(G[0] + sum_n mu^{n+1} sig^n) / k``,
    ``dJ/dcn_m = G[m]``, ``dJ/dcw_m = G[K+m]``, carried to mass and stiffness through the
    coefficient map; zero off the undamped interior.  ``sum_shots``: the sum over shots
    ``(1, n0, n1)``, by one pass through the (linear) chain rule on the shot-summed sums."""
    g = med.grid
    K = med.K
    n0, n1 = g.shape
    Gm = G[0][:, :n0, :n1].clone()
    if mu_src is not None:
        s_sig = solver._sig_tensor(g, src, sig, mu_src.shape[1])
        contrib = (mu_src.double() * s_sig.double()[:, :, : src.count]).sum(1)
        for s, f in enumerate(src.nodes_flat):
            if len(f):
                Gm[s].reshape(-1).index_add_(0, torch.as_tensor(f, device=Gm.device), contrib[s, : len(f)])
    if sum_shots:
        Gm = Gm.sum(0, keepdim=True)
        G = G.sum(1, keepdim=True)
    und = torch.as_tensor(med.undamped, device=Gm.device)
    dm, ds = [], []
    with torch.enable_grad():  # also inside an autograd.Function's backward
        m = med.mass.detach().clone().requires_grad_(True)
        st = med.stiff.detach().clone().requires_grad_(True)
        k, cn, cw = face_operator(m, st, g.spacing, g.dt, med.closure, g.order)
        kd = k.detach()
        for s in range(Gm.shape[0]):
            dk = torch.where(und, -Gm[s] / torch.where(und, kd, torch.ones_like(kd)), torch.zeros_like(kd))
            gn = torch.stack([G[a][s] for a in range(1, K + 1)])
            gw = torch.stack([G[a][s] for a in range(K + 1, 2 * K + 1)])
            gm_, gs_ = torch.autograd.grad([k, cn, cw], [m, st], [dk, gn, gw], retain_graph=True)
            dm.append(torch.where(und, gm_, torch.zeros_like(gm_)))
            ds.append(torch.where(und, gs_, torch.zeros_like(gs_)))
    return torch.stack(dm), torch.stack(ds)


class _AcTraces(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mass, stiff, grid, closure, beta_profiles, src, sig, rec):
        med = AcMedium(grid, mass.detach(), stiff.detach(), closure, beta_profiles)
        strip = PointSet(grid, [tuple(int(c) for c in p) for p in med.strip_nodes()])
        nsteps = int(sig.shape[-2])
        traces, store, state = forward(grid, med, nsteps, src, sig, rec, strip)
        ctx.keep = (grid, med, src, sig, rec, strip, store, state, nsteps)
        return traces

    @staticmethod
    def backward(ctx, grad_traces):
        grid, med, src, sig, rec, strip, store, state, nsteps = ctx.keep
        G, mu_src = backward(grid, med, nsteps, state, src, sig, rec, grad_traces.contiguous(), strip, store)
        dm, ds = material_gradient(med, G, mu_src, sig, src, sum_shots=True)
        ctx.keep = None
        return dm[0], ds[0], None, None, None, None, None, None


def acoustic_traces(mass, stiff, grid: WaveGrid, src: Optional[PointSet], sig, rec: PointSet,
                    closure: str = "neumann", beta_profiles=None) -> torch.Tensor:
    """Receiver traces ``(nshots, nsteps, R)`` of the variable mass / stiffness model, differentiable
    in ``mass`` and ``stiff`` (node arrays) through the hand-written reconstruction adjoint."""
    sig = torch.as_tensor(sig, device=grid.device)
    return _AcTraces.apply(torch.as_tensor(mass, dtype=torch.float64, device=grid.device),
                           torch.as_tensor(stiff, dtype=torch.float64, device=grid.device),
                           grid, closure, beta_profiles, src, sig, rec)


# --------------------------------------------------------------------- cuwave facade
@dataclass(frozen=True)
class AcousticWave:
    """Simulation descriptor of cuwave's ``AcousticWave`` (fields as at
    ``topopt_acoustic2D.py:119-131``); node conventions as ``ScalarWave``.  ``damping`` takes the
    per-axis profiles of :func:`neuralmech_amd.boundary.sponge`."""

    Nx: Sequence[int]
    dx: Sequence[float]
    N: int
    dt: float
    threads: Optional[Sequence[int]] = None
    precision: str = "float32"
    space_order: int = 2
    rho1: float = 1.0
    rho2: float = 1.0
    kappa1: float = 1.0
    kappa2: float = 1.0
    damping: Optional[tuple] = None
    boundary: str = "neumann"
    device: str = "cuda"
    _cache: dict = field(default_factory=dict, compare=False, repr=False)

    def __post_init__(self):
        object.__setattr__(self, "Nx", tuple(int(n) for n in self.Nx))
        dx = self.dx if np.ndim(self.dx) else [self.dx] * len(self.Nx)
        object.__setattr__(self, "dx", tuple(float(h) for h in dx))
        if len(self.Nx) != 2 or len(self.dx) != 2:
            raise ValueError("AcousticWave is two-dimensional")
        if self.precision not in _PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}")
        if self.space_order not in (2, 4, 6, 8):
            raise ValueError("space_order must be 2, 4, 6 or 8")
        if self.boundary not in CLOSURES:
            raise ValueError(f"boundary must be one of {CLOSURES}")
        for name in ("rho1", "rho2", "kappa1", "kappa2"):
            if not float(getattr(self, name)) > 0:
                raise ValueError(f"{name} must be > 0")
        limit = solver.stable_dt(self.dx, self.max_wavespeed, self.space_order)
        if self.dt > limit * (1 + 1e-12):
            raise ValueError(f"dt = {self.dt:.6g} exceeds the stable step {limit:.6g} for wave speed "
                             f"{self.max_wavespeed:.6g}")

    @property
    def max_wavespeed(self) -> float:
        return float(max(np.sqrt(self.kappa1 / self.rho1), np.sqrt(self.kappa2 / self.rho2)))

    @property
    def ndim(self) -> int:
        return 2

    @property
    def dtype(self) -> torch.dtype:
        return _PRECISIONS[self.precision]

    @property
    def Nx_padded(self) -> tuple:
        return self.Nx

    @property
    def cell_volume(self) -> float:
        return float(np.prod(self.dx))

    def _gamma(self, gamma) -> torch.Tensor:
        g = torch.as_tensor(gamma if gamma is not None else 1.0, dtype=torch.float64).to(self.device)
        g = g.expand(self.Nx) if g.dim() == 0 else g
        if tuple(g.shape) != self.Nx:
            raise ValueError(f"indicator must have the node shape {self.Nx}")
        return g

    def material(self, gamma):
        """``(mass, stiffness)`` node arrays: ``1/kappa`` and ``1/rho`` interpolated linearly."""
        g = self._gamma(gamma)
        m = (1.0 - g) / self.kappa1 + g / self.kappa2
        s = (1.0 - g) / self.rho1 + g / self.rho2
        return m, s

    def parametrization_jacobian(self, gamma):
        """``(d mass / d gamma, d stiffness / d gamma)`` on the nodes."""
        g = self._gamma(gamma)
        return (torch.full_like(g, 1.0 / self.kappa2 - 1.0 / self.kappa1),
                torch.full_like(g, 1.0 / self.rho2 - 1.0 / self.rho1))

    def grid(self, nshots: int = 1) -> WaveGrid:
        return WaveGrid(self.Nx, self.dx, self.dt, self.space_order, self.dtype, nshots=nshots, device=self.device)


def _node_list(nodes):
    """Node indices as a list of (i, j) tuples from a (D, S) array, a (S, D) array or a Sensors."""
    if hasattr(nodes, "nodes"):
        nodes = nodes.nodes
    a = np.asarray(nodes.cpu() if torch.is_tensor(nodes) else nodes, dtype=np.int64)
    if a.ndim == 1:
        a = a.reshape(2, -1)
    if a.shape[0] == 2 and a.ndim == 2:
        a = a.T
    return [tuple(int(c) for c in p) for p in a]


def _scale_of(sim):
    """Per-node factor on source signals (``ScalarWave`` with an indicator: ``rho c^2``)."""
    f = getattr(sim, "source_scale", None)
    return None if f is None else np.asarray(f().detach().cpu(), np.float64)


def _shots(grid, sources, nsteps, scale=None):
    """PointSet (one list per shot) and ``(nshots, nsteps, width)`` signal of ``Source`` objects."""
    nodes = [[tuple(int(c) for c in p) for p in src.idx.T] for src in sources]
    pts = PointSet(grid, nodes, per_shot=True) if len(sources) > 1 else PointSet(grid, nodes[0])
    width = max(src.signal.shape[1] for src in sources)
    sig = np.zeros((len(sources), nsteps, width))
    for i, src in enumerate(sources):
        n = min(nsteps, src.signal.shape[0])
        col = src.signal[:n]
        if scale is not None:
            col = col * scale[tuple(src.idx)][None, :]
        sig[i, :n, : src.signal.shape[1]] = col
    return pts, torch.as_tensor(sig)


def _source_setup(grid, source, nsteps, scale=None):
    """PointSet and ``(1, nsteps, S)`` signal of one ``Source`` (or None)."""
    if isinstance(source, (list, tuple)):
        if len(source) != 1:
            raise ValueError("one Source per call (a list of sources is a batch of shots)")
        source = source[0]
    if source is None:
        return None, None
    return _shots(grid, [source], nsteps, scale)


def _material(sim, gamma):
    if len(sim.Nx) != 2:
        raise NotImplementedError("the density indicator is implemented for 2D grids")
    return sim.material(gamma)


def _closure(sim):
    return getattr(sim, "boundary", "neumann") if isinstance(sim, AcousticWave) else "dirichlet"


def _medium(sim, gamma, nshots=1):
    grid = sim.grid(nshots=nshots)
    m, s = _material(sim, gamma)
    return grid, AcMedium(grid, m, s, _closure(sim), sim.damping)


def measure_indicator(sim, sources, gamma, sensors) -> list:
    """Traces ``(N, R)`` of every source, all shots in one launch per step."""
    sources = list(sources)
    grid, med = _medium(sim, gamma, len(sources))
    src, sig = _shots(grid, sources, sim.N, _scale_of(sim))
    rec = PointSet(grid, _node_list(sensors))
    tr, _, _ = forward(grid, med, sim.N, src, sig, rec)
    return [tr[k] for k in range(len(sources))]


def shots_misfit(sim, sources, gamma, sensors, observed, gradient: bool = True):
    """``1/2 sum (d - d_obs)^2`` over shots, and with ``gradient`` its derivative in the indicator:
    one batched forward (strip recorded) and one reconstruction adjoint over all shots."""
    sources = list(sources)
    grid, med = _medium(sim, gamma, len(sources))
    src, sig = _shots(grid, sources, sim.N, _scale_of(sim))
    rec = PointSet(grid, _node_list(sensors))
    sn = med.strip_nodes() if gradient else np.zeros((0, 2), np.int64)
    strip = PointSet(grid, [tuple(int(c) for c in p) for p in sn]) if len(sn) else None
    tr, store, state = forward(grid, med, sim.N, src, sig, rec, strip)
    obs = torch.as_tensor(observed, dtype=torch.float64).to(grid.device)
    if obs.shape != tr.shape:
        raise ValueError(f"observed records must be (shots, N, receivers) = {tuple(tr.shape)}")
    r = tr.double() - obs
    cost = 0.5 * float((r * r).sum())
    if not gradient:
        return cost
    G, mu_src = backward(grid, med, sim.N, state, src, sig, rec, r, strip, store)
    dm, ds = material_gradient(med, G, mu_src, sig, src, sum_shots=True)
    d_mass, d_stiff = sim.parametrization_jacobian(gamma)
    return cost, d_mass * dm[0] + d_stiff * ds[0]


def simulate_acoustic(sim, source=None, indicator=None, sensors=None, record_every=None):
    """``simulate`` for a variable mass / stiffness descriptor (see ``wave.simulate``)."""
    grid, med = _medium(sim, indicator)
    src, sig = _source_setup(grid, source, sim.N, _scale_of(sim))
    rec = PointSet(grid, _node_list(sensors)) if sensors is not None else None
    state = solver.WaveState.rest(grid)
    frames, traces, step = [], [], 0
    chunk = int(record_every) if record_every else sim.N
    while step < sim.N:
        n = min(chunk, sim.N - step)
        tr, _, state = forward(grid, med, n, src, sig, rec, state=state)
        if tr is not None:
            traces.append(tr[0])
        step += n
        if record_every:
            frames.append(grid.nodes(state[1])[0].detach().cpu().numpy().copy())
    u = grid.nodes(state[1])[0].clone()
    out = [u]
    if sensors is not None:
        out.append(torch.cat(traces, 0) if traces else torch.zeros(sim.N, 0, dtype=sim.dtype, device=sim.device))
    if record_every:
        out.append(frames)
    return out[0] if len(out) == 1 else tuple(out)


def energy(sim):
    """Objective ``J = dt sum_t sum_r u^2`` of sensor traces, with ``dJ/du = 2 dt u``."""
    dt = float(sim.dt)

    def objective(traces):
        t = torch.as_tensor(traces)
        return float((t.double() ** 2).sum()) * dt, 2.0 * dt * t.double()

    return objective


def response(sim, source, gamma, sensors, objective):
    """``(cost, traces)``: one forward run and the objective on the sensor traces ``(N, S)``."""
    grid, med = _medium(sim, gamma)
    src, sig = _source_setup(grid, source, sim.N, _scale_of(sim))
    rec = PointSet(grid, _node_list(sensors))
    tr, _, _ = forward(grid, med, sim.N, src, sig, rec)
    cost, _ = objective(tr[0])
    return cost, tr[0]


def reconstruction_nodes(sim):
    """``(strip, undamped)``: the strip of nodes the forward records for the reconstruction,
    ``(2, S)`` node indices, and the mask of nodes where gradients are available."""
    prof = None if sim.damping is None else [np.asarray(torch.as_tensor(p, dtype=torch.float64).cpu())
                                               for p in sim.damping]
    npdt = np.float32 if sim.precision == "float32" else np.float64
    und = np.zeros(sim.Nx, bool)
    if prof is None:
        und[1:-1, 1:-1] = True
    else:
        p0, p1 = (np.asarray(p, np.float64).astype(npdt) for p in prof)
        und[1:-1, 1:-1] = (p0[:, None] + p1[None, :])[1:-1, 1:-1] == 0
    return _strip(und, sim.space_order // 2).T.copy(), und


def reconstruction_sensitivity(sim, source, gamma, nodes, objective):
    """``(loss, {"mass", "stiff"}, traces, p_N)``: forward with the strip recorded, the objective on
    the traces at ``nodes`` (``(2, S)``) and the reconstruction adjoint; gradients with respect to
    the mass and stiffness node arrays (zero where the sponge damps)."""
    grid, med = _medium(sim, gamma)
    src, sig = _source_setup(grid, source, sim.N, _scale_of(sim))
    rec = PointSet(grid, _node_list(nodes))
    strip = PointSet(grid, [tuple(int(c) for c in p) for p in med.strip_nodes()])
    tr, store, state = forward(grid, med, sim.N, src, sig, rec, strip)
    traces = tr[0]
    loss, dJ = objective(traces)
    p_last = grid.nodes(state[1])[0].clone()
    G, mu_src = backward(grid, med, sim.N, state, src, sig, rec, torch.as_tensor(dJ)[None], strip, store)
    dm, ds = material_gradient(med, G, mu_src, sig, src)
    return loss, {"mass": dm[0], "stiff": ds[0]}, traces, p_last


def response_gradient(sim, source, gamma, sensors, objective, adjoint=reconstruction_sensitivity):
    """``(cost, dcost/dgamma)`` through ``adjoint`` (``reconstruction_sensitivity``) and the
    indicator parametrisation."""
    nodes = sensors.nodes if hasattr(sensors, "nodes") else sensors
    loss, grads, _, _ = adjoint(sim, source, gamma, nodes, objective)
    d_mass, d_stiff = sim.parametrization_jacobian(gamma)
    return loss, d_mass * grads["mass"] + d_stiff * grads["stiff"]


__all__ = ["AcMedium", "AcousticWave", "acoustic_traces", "backward", "energy", "face_operator", "forward",
           "material_gradient", "measure_indicator", "reconstruction_nodes", "reconstruction_sensitivity",
           "response", "response_gradient", "shots_misfit", "simulate_acoustic", "volume_weights"]
