"""ctypes front end of the C oracle (``liboracle_wavefd.so``, built by ``oracle/Makefile``).

TEST INFRASTRUCTURE ONLY: used by ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` as the checker, never by the product package.

All arrays are logical C-order node arrays ``(n0[, n1[, n2]])`` including the boundary ring.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_DIR = Path(__file__).resolve().parent
_LIB_PATH = _DIR / "liboracle_wavefd.so"
_lib = None

_i64p = ctypes.POINTER(ctypes.c_int64)
_dp = ctypes.POINTER(ctypes.c_double)


def build() -> Path:
    """Compile the C oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", str(_DIR), "-B", "liboracle_wavefd.so"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(_LIB_PATH))
        for sfx, ct in (("f32", ctypes.c_float), ("f64", ctypes.c_double)):
            tp = ctypes.POINTER(ct)
            fwd = getattr(_lib, f"wfo_forward_{sfx}")
            fwd.restype = ctypes.c_int
            fwd.argtypes = [ctypes.c_int, _i64p, _dp, ctypes.c_int, tp, ctypes.c_double, tp,
                            ctypes.c_int, _i64p, tp, ctypes.c_int, _i64p, tp, tp, tp,
                            ctypes.c_int64, ctypes.c_int]
            adj = getattr(_lib, f"wfo_adjoint_{sfx}")
            adj.restype = ctypes.c_int
            adj.argtypes = [ctypes.c_int, _i64p, _dp, ctypes.c_int, tp, ctypes.c_double, tp,
                            ctypes.c_int, _i64p, tp, ctypes.c_int, _i64p, tp, tp,
                            ctypes.c_int64, ctypes.c_int64, _dp, ctypes.c_int]
            adjs = getattr(_lib, f"wfo_adjoint_shots_{sfx}")
            adjs.restype = ctypes.c_int
            adjs.argtypes = [ctypes.c_int, _i64p, _dp, ctypes.c_int, tp, ctypes.c_double, tp,
                             ctypes.c_int, ctypes.c_int, ctypes.c_int, _i64p, ctypes.c_int, tp,
                             ctypes.c_int, _i64p, ctypes.c_int, tp, tp,
                             ctypes.c_int64, ctypes.c_int64, _dp, tp, ctypes.c_int]
            lap = getattr(_lib, f"wfo_lap_{sfx}")
            lap.restype = ctypes.c_int
            lap.argtypes = [ctypes.c_int, _i64p, _dp, ctypes.c_int, tp, tp]
            w = getattr(_lib, f"wfo_weights_{sfx}")
            w.restype = ctypes.c_int
            w.argtypes = [ctypes.c_int, _dp, ctypes.c_int, tp, tp]
        _lib.wfo_fd_coeffs.restype = ctypes.c_int
        _lib.wfo_fd_coeffs.argtypes = [ctypes.c_int, _dp]
    return _lib


def _sfx(dtype) -> str:
    dtype = np.dtype(dtype)
    if dtype == np.float32:
        return "f32"
    if dtype == np.float64:
        return "f64"
    raise TypeError(f"oracle dtype must be float32 or float64, got {dtype}")


def _ptr(a, ct):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ct))


def _ct(dtype):
    return ctypes.c_float if np.dtype(dtype) == np.float32 else ctypes.c_double


def fd_coeffs(order: int) -> np.ndarray:
    out = np.zeros(5)
    if lib().wfo_fd_coeffs(order, _ptr(out, ctypes.c_double)):
        raise ValueError(order)
    return out[: order // 2 + 1]


def weights(ndim: int, h, order: int, dtype=np.float32):
    """``(wc, w[ndim, K])`` exactly as the oracle's step uses them."""
    ct = _ct(dtype)
    h = np.ascontiguousarray(h, np.float64)
    wc = np.zeros(1, dtype)
    w = np.zeros(12, dtype)
    if getattr(lib(), f"wfo_weights_{_sfx(dtype)}")(ndim, _ptr(h, ctypes.c_double), order,
                                                       _ptr(wc, ct), _ptr(w, ct)):
        raise ValueError((ndim, order))
    K = order // 2
    return wc[0], w[: ndim * K].reshape(ndim, K)


def _nodes(nodes, shape) -> np.ndarray:
    """Flat logical node indices from index tuples or flat ints."""
    if len(nodes) == 0:
        return np.zeros(1, np.int64)
    arr = np.asarray(nodes)
    if arr.ndim == 2:
        return np.ravel_multi_index(tuple(arr.T), shape).astype(np.int64)
    return arr.astype(np.int64)


def forward(kc, order, h, nsteps, *, kconst=0.0, beta_profiles=None, src_nodes=(), src_sig=None,
            rec_nodes=(), u_prev=None, u_cur=None, dtype=None, shape=None, nthreads=None):
    """Forward run; returns ``(traces[nsteps, R], u_prev, u_cur)``.

    ``kc`` is the per-node coefficient array (or ``None`` with ``kconst`` and ``shape``);
    ``beta_profiles`` a list of per-axis damping profiles (slowest axis first).
    """
    if kc is not None:
        kc = np.ascontiguousarray(kc)
        dtype = kc.dtype if dtype is None else np.dtype(dtype)
        kc = kc.astype(dtype, copy=False)
        shape = kc.shape
    dtype = np.dtype(dtype)
    ct = _ct(dtype)
    ndim = len(shape)
    n = np.asarray(shape, np.int64)
    h = np.ascontiguousarray(h, np.float64)
    up = np.zeros(shape, dtype) if u_prev is None else np.ascontiguousarray(u_prev, dtype).copy()
    uc = np.zeros(shape, dtype) if u_cur is None else np.ascontiguousarray(u_cur, dtype).copy()
    beta = None
    if beta_profiles is not None:
        beta = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype) for p in beta_profiles]))
    src = _nodes(src_nodes, shape)
    rec = _nodes(rec_nodes, shape)
    nsrc, nrec = len(src_nodes), len(rec_nodes)
    sig = None if src_sig is None else np.ascontiguousarray(src_sig, dtype)
    traces = np.zeros((nsteps, max(nrec, 1)), dtype)
    nthreads = nthreads or os.cpu_count() or 1
    rc = getattr(lib(), f"wfo_forward_{_sfx(dtype)}")(
        ndim, _ptr(n, ctypes.c_int64), _ptr(h, ctypes.c_double), order, _ptr(kc, ct), float(kconst),
        _ptr(beta, ct), nsrc, _ptr(src, ctypes.c_int64), _ptr(sig, ct), nrec,
        _ptr(rec, ctypes.c_int64), _ptr(traces, ct), _ptr(up, ct), _ptr(uc, ct), nsteps, nthreads)
    if rc:
        raise RuntimeError(f"oracle forward failed ({rc})")
    return traces[:, :nrec], up, uc


def adjoint(kc, order, h, nsteps, adj_sig, u_hist, *, seg, kconst=0.0, beta_profiles=None,
            rec_nodes=(), src_nodes=(), dtype=None, shape=None, nthreads=None):
    """Adjoint sweep + imaging; returns ``(G[shape] float64, mu_src[nsteps, S])``.

    ``G = sum_n mu^{n+1} L u^n`` with ``T`` partials flushed into float64 whenever
    ``(m-1) % seg == 0``; ``adj_sig[n, r]`` is injected at receiver ``r`` at adjoint step
    ``m = n + 1``; ``u_hist[n] = u^n``.
    """
    if kc is not None:
        kc = np.ascontiguousarray(kc)
        dtype = kc.dtype if dtype is None else np.dtype(dtype)
        kc = kc.astype(dtype, copy=False)
        shape = kc.shape
    dtype = np.dtype(dtype)
    ct = _ct(dtype)
    n = np.asarray(shape, np.int64)
    h = np.ascontiguousarray(h, np.float64)
    beta = None
    if beta_profiles is not None:
        beta = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype) for p in beta_profiles]))
    rec = _nodes(rec_nodes, shape)
    src = _nodes(src_nodes, shape)
    sig = np.ascontiguousarray(adj_sig, dtype)
    hist = np.ascontiguousarray(u_hist, dtype)
    mu_src = np.zeros((nsteps, max(len(src_nodes), 1)), dtype)
    G = np.zeros(shape, np.float64)
    nthreads = nthreads or os.cpu_count() or 1
    rc = getattr(lib(), f"wfo_adjoint_{_sfx(dtype)}")(
        len(shape), _ptr(n, ctypes.c_int64), _ptr(h, ctypes.c_double), order, _ptr(kc, ct),
        float(kconst), _ptr(beta, ct), len(rec_nodes), _ptr(rec, ctypes.c_int64), _ptr(sig, ct),
        len(src_nodes), _ptr(src, ctypes.c_int64), _ptr(mu_src, ct), _ptr(hist, ct), nsteps, seg,
        _ptr(G, ctypes.c_double), nthreads)
    if rc:
        raise RuntimeError(f"oracle adjoint failed ({rc})")
    return G, mu_src[:, : len(src_nodes)]


def adjoint_shots(kc, order, h, nsteps, adj_sig, u_hist, *, seg, group, beta_profiles=None,
                  rec_nodes=(), src_nodes=(), keep_mu=False, nthreads=None):
    """Adjoint sweeps of ``S`` shots over one medium; returns ``(G float64, mu_src[S, nsteps, P])``,
    plus ``mu[S, nsteps, ...]`` (``mu[s, m-1] = mu_s^m``) when ``keep_mu``.

    ``adj_sig[s, n, r]`` and ``u_hist[s, n] = u_s^n`` per shot.  ``src_nodes`` / ``rec_nodes`` are
    one node list for every shot, or ``S`` lists (one per shot).  The imaging partials are one per
    group of ``group`` consecutive shots, each summing its shots' terms in shot order at every
    step, flushed into ``G`` in group order whenever ``(m-1) % seg == 0``.
    """
    kc = np.ascontiguousarray(kc)
    dtype = kc.dtype
    ct = _ct(dtype)
    shape = kc.shape
    sig = np.ascontiguousarray(adj_sig, dtype)
    hist = np.ascontiguousarray(u_hist, dtype)
    nshots = sig.shape[0]
    assert sig.shape[1] == nsteps and hist.shape == (nshots, nsteps) + shape

    def node_list(nodes):
        per_shot = len(nodes) == nshots and len(nodes) > 0 and isinstance(nodes[0], (list, tuple)) \
            and len(nodes[0]) > 0 and isinstance(nodes[0][0], (list, tuple))
        lists = list(nodes) if per_shot else [nodes]
        cnt = len(lists[0])
        flat = np.concatenate([_nodes(l, shape)[: max(cnt, 1)] for l in lists]) if cnt else np.zeros(1, np.int64)
        return np.ascontiguousarray(flat, np.int64), cnt, int(per_shot)

    src, nsrc, src_ps = node_list(src_nodes)
    rec, nrec, rec_ps = node_list(rec_nodes)
    assert sig.shape[2] == nrec
    beta = None
    if beta_profiles is not None:
        beta = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype) for p in beta_profiles]))
    mu_src = np.zeros((nshots, nsteps, max(nsrc, 1)), dtype)
    G = np.zeros(shape, np.float64)
    mu = np.zeros((nshots, nsteps) + shape, dtype) if keep_mu else None
    nthreads = nthreads or os.cpu_count() or 1
    rc = getattr(lib(), f"wfo_adjoint_shots_{_sfx(dtype)}")(
        len(shape), _ptr(np.asarray(shape, np.int64), ctypes.c_int64),
        _ptr(np.ascontiguousarray(h, np.float64), ctypes.c_double), order, _ptr(kc, ct), 0.0,
        _ptr(beta, ct), nshots, group, nrec, _ptr(rec, ctypes.c_int64), rec_ps, _ptr(sig, ct),
        nsrc, _ptr(src, ctypes.c_int64), src_ps, _ptr(mu_src, ct), _ptr(hist, ct), nsteps, seg,
        _ptr(G, ctypes.c_double), _ptr(mu, ct), nthreads)
    if rc:
        raise RuntimeError(f"oracle adjoint failed ({rc})")
    return (G, mu_src[:, :, :nsrc], mu) if keep_mu else (G, mu_src[:, :, :nsrc])


def lap(u, order, h):
    """``L u`` at the interior nodes exactly as the oracle's imaging sum computes it."""
    u = np.ascontiguousarray(u)
    out = np.zeros_like(u)
    ct = _ct(u.dtype)
    rc = getattr(lib(), f"wfo_lap_{_sfx(u.dtype)}")(
        u.ndim, _ptr(np.asarray(u.shape, np.int64), ctypes.c_int64),
        _ptr(np.ascontiguousarray(h, np.float64), ctypes.c_double), order, _ptr(u, ct), _ptr(out, ct))
    if rc:
        raise RuntimeError(f"oracle lap failed ({rc})")
    return out


def history(kc, order, h, nsteps, **kw):
    """Forward fields ``u^0..u^{nsteps-1}`` (stacked) plus traces, one oracle step at a time."""
    kc = np.ascontiguousarray(kc)
    shape = kc.shape
    dtype = kc.dtype
    up = np.zeros(shape, dtype)
    uc = np.zeros(shape, dtype)
    src_sig = kw.pop("src_sig", None)
    hist = np.zeros((nsteps,) + shape, dtype)
    traces = []
    for s in range(nsteps):
        hist[s] = uc
        sig = None if src_sig is None else src_sig[s : s + 1]
        tr, up, uc = forward(kc, order, h, 1, u_prev=up, u_cur=uc, src_sig=sig, **kw)
        traces.append(tr[0])
    return np.stack(traces), hist, up, uc
