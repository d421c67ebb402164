"""Seeded test problems shared by the CPU and GPU tests, and runners for the oracle and the HIP
library on the same inputs."""

from __future__ import annotations

import math

import numpy as np

from oracle import c_oracle, slice_ref


def make_case(shape, order, *, seed=0, varc=True, sponge=True, nsrc=2, nrec=6, nsteps=40, safety=0.6,
              h=None):
    """A small random medium with sources, receivers and a sponge (float64 description)."""
    rng = np.random.default_rng(seed)
    ndim = len(shape)
    h = tuple(h) if h is not None else tuple(0.8 + 0.1 * d for d in range(ndim))
    v = rng.uniform(1.0, 2.0, shape) if varc else None
    vmax = 2.0 if varc else 1.5
    dt = safety * slice_ref.stable_dt(h, vmax, order)
    kconst = (1.5 * dt) ** 2
    kc = (v * dt) ** 2 if varc else None
    interior = [rng.integers(1, n - 1, size=max(nsrc, nrec) * 2) for n in shape]
    pts = list(zip(*interior))
    src = [tuple(int(x) for x in p) for p in pts[:nsrc]]
    rec = [tuple(int(x) for x in p) for p in pts[nsrc:nsrc + nrec]]
    sig = rng.standard_normal((nsteps, nsrc))
    prof = None
    if sponge:
        prof = [slice_ref.sponge_profile(n, max(1, n // 5), 0.15) for n in shape]
    return dict(shape=tuple(shape), order=order, h=h, dt=dt, v=v, kc=kc, kconst=kconst, src=src, rec=rec,
                sig=sig, prof=prof, nsteps=nsteps)


def oracle_forward(case, dtype):
    kc = None if case["kc"] is None else case["kc"].astype(dtype)
    kconst = float(np.asarray(case["kconst"], dtype))
    tr, up, uc = c_oracle.forward(kc, case["order"], case["h"], case["nsteps"], kconst=kconst,
                                  beta_profiles=case["prof"], src_nodes=case["src"],
                                  src_sig=case["sig"].astype(dtype) if case["src"] else None,
                                  rec_nodes=case["rec"], dtype=dtype, shape=case["shape"])
    return tr, up, uc


def gpu_objects(case, torch_dtype, nshots=1):
    import torch

    from neuralmech_amd import solver

    g = solver.WaveGrid(case["shape"], case["h"], case["dt"], case["order"], torch_dtype, nshots=nshots)
    if case["kc"] is not None:
        kc = torch.as_tensor(case["kc"]).to(torch_dtype).double()  # the oracle rounds kc the same way
        med = solver.Medium(g, kc=kc, beta_profiles=_prof(case, torch_dtype))
    else:
        kconst = float(np.asarray(case["kconst"], np.float32 if torch_dtype == torch.float32 else np.float64))
        med = solver.Medium(g, kconst=kconst, beta_profiles=_prof(case, torch_dtype))
    src = solver.PointSet(g, case["src"]) if case["src"] else None
    rec = solver.PointSet(g, case["rec"]) if case["rec"] else None
    return g, med, src, rec


def _prof(case, torch_dtype):
    if case["prof"] is None:
        return None
    import torch

    return [torch.as_tensor(p).to(torch_dtype).double() for p in case["prof"]]


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def ricker_256_order2_setup():
    """256 x 256 nodes, dx = 1, c = 1, c dt / dx = 0.5, order 2, Ricker f0 = 0.05/dt at the centre,
    64 receivers on row 20 (the reference's own CPU-runnable case)."""
    n = 256
    dt = 0.5
    f0 = 0.05 / dt
    nsteps = 1000
    t = np.arange(nsteps) * dt
    sig = slice_ref.ricker(t, f0) / (1.0 * 1.0)  # / cell volume dx*dy
    rec = [(20, int(j)) for j in np.round(np.linspace(1, n - 2, 64)).astype(int)]
    return dict(shape=(n, n), order=2, h=(1.0, 1.0), dt=dt, v=None, kc=None, kconst=(1.0 * dt) ** 2,
                src=[(n // 2, n // 2)], rec=rec, sig=sig[:, None], prof=None, nsteps=nsteps)


def ceil_div(a, b):
    return -(-a // b)


def n_interior(shape):
    return math.prod(s - 2 for s in shape)
