"""Pins of the CPU oracle by what the mathematics fixes: closed-form weights and stability
limits, plane-wave symbols and exact discrete plane waves, Dirichlet eigenmodes, the 1D magic
time step, discrete energy, reciprocity, linearity, the adjoint against torch autograd, and
mutation checks that plant a mistake in the oracle and watch a pin fail."""

from fractions import Fraction as F
import math

import numpy as np
import pytest

from oracle import c_oracle, slice_ref, torch_ref
from tests.wave_cases import make_case, rel_l2

TABLE = {  # textbook central weights a_0..a_K
    2: [F(-2), F(1)],
    4: [F(-5, 2), F(4, 3), F(-1, 12)],
    6: [F(-49, 18), F(3, 2), F(-3, 20), F(1, 90)],
    8: [F(-205, 72), F(8, 5), F(-1, 5), F(8, 315), F(-1, 560)],
}


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_weights_closed_form(order):
    a = slice_ref.fd_weights(order)
    assert a == TABLE[order]
    # moment conditions: exact second derivative of x^2, annihilates x^{2j} for 2 <= j <= K
    K = order // 2
    assert a[0] + 2 * sum(a[1:]) == 0
    assert 2 * sum(a[k] * k ** 2 for k in range(1, K + 1)) == 2
    for j in range(2, K + 1):
        assert sum(a[k] * k ** (2 * j) for k in range(1, K + 1)) == 0
    # the C oracle rounds each rational once
    assert np.array_equal(c_oracle.fd_coeffs(order), np.array([float(x) for x in a]))


@pytest.mark.parametrize("order,s,lim2,lim3", [(2, F(4), 0.7071, 0.5774), (4, F(16, 3), 0.6124, 0.5000),
                                               (6, F(272, 45), 0.5752, 0.4697), (8, F(2048, 315), 0.5546, 0.4529)])
def test_stability_limits(order, s, lim2, lim3):
    assert slice_ref.stability_sum(order) == s
    assert round(slice_ref.stable_dt((1.0, 1.0), 1.0, order), 4) == lim2
    assert round(slice_ref.stable_dt((1.0, 1.0, 1.0), 1.0, order), 4) == lim3


def test_order2_stable_dt_is_the_reference_formula():
    dx = 1.0 / 1599
    assert slice_ref.stable_dt((dx, dx), 1.0, 2) == pytest.approx(slice_ref.reference_dt(dx, dx, 1.0), rel=1e-15)
    # unequal spacing: the reference's min(dx, dy)/c/sqrt(2) is the conservative side
    dx, dy = 1.0 / 1599, 2.0 / 3199
    assert slice_ref.stable_dt((dx, dy), 1.0, 2) >= slice_ref.reference_dt(dx, dy, 1.0)


def _symbol(order, xi, h):
    a = [float(x) for x in slice_ref.fd_weights(order)]
    return -(a[0] + 2 * sum(a[k] * math.cos(k * xi * h) for k in range(1, len(a)))) / h ** 2


@pytest.mark.parametrize("order", [2, 4, 6, 8])
@pytest.mark.parametrize("impl", ["slice", "c"])
def test_laplacian_plane_wave_symbol(order, impl):
    """L cos(xi.x) = -(sum_d lambda_d(xi_d)) cos(xi.x) at nodes K away from the boundary."""
    shape, h = (40, 47), (0.7, 1.3)
    xi = (0.9, 0.4)
    x = np.meshgrid(*[np.arange(n) * hd for n, hd in zip(shape, h)], indexing="ij")
    u = np.cos(xi[0] * x[0] + xi[1] * x[1] + 0.3)
    lam = sum(_symbol(order, k, hd) for k, hd in zip(xi, h))
    K = order // 2
    inner = tuple(slice(K + 1, n - K - 1) for n in shape)
    if impl == "slice":
        lap = slice_ref.laplacian(u, order, h)
    else:
        # one step from (u_prev = 0, u_cur = u) with kc = 1: u_next = 2u + L u
        kc = np.ones(shape)
        _, _, un = c_oracle.forward(kc, order, h, 1, u_prev=np.zeros(shape), u_cur=u)
        lap = un - 2 * u
    assert np.max(np.abs(lap[inner] + lam * u[inner])) < 1e-11 * max(1.0, abs(lam))


@pytest.mark.parametrize("order", [4, 8])
def test_exact_discrete_plane_wave_in_time(order):
    """cos(xi.x - w t) with sin^2(w dt/2) = (c dt/2)^2 sum lambda_d is an exact discrete solution away
    from the region the boundary can influence in n steps."""
    shape, h, c = (150, 140), (1.0, 1.0), 1.0
    xi = (0.5, 0.3)
    dt = 0.5 * slice_ref.stable_dt(h, c, order)
    lam = sum(_symbol(order, k, hd) for k, hd in zip(xi, h))
    w = 2 * math.asin(c * dt / 2 * math.sqrt(lam)) / dt
    x = np.meshgrid(*[np.arange(n) * hd for n, hd in zip(shape, h)], indexing="ij")
    ph = xi[0] * x[0] + xi[1] * x[1]
    kc = np.full(shape, (c * dt) ** 2)
    n = 12
    _, _, un = c_oracle.forward(kc, order, h, n, u_prev=np.cos(ph + w * dt), u_cur=np.cos(ph))
    m = (n + 1) * (order // 2) + 1
    inner = (slice(m, shape[0] - m), slice(m, shape[1] - m))
    assert np.max(np.abs(un[inner] - np.cos(ph - w * n * dt)[inner])) < 1e-10


def test_dirichlet_eigenmode_order2():
    """sin(p pi i/(N-1)) sin(q pi j/(M-1)) evolves as cos(n theta), cos theta = 1 - c^2 dt^2 lambda / 2."""
    N, M, p, q = 33, 45, 3, 5
    h = (0.9, 1.1)
    dt = 0.8 * slice_ref.stable_dt(h, 1.0, 2)
    i, j = np.meshgrid(np.arange(N), np.arange(M), indexing="ij")
    mode = np.sin(p * math.pi * i / (N - 1)) * np.sin(q * math.pi * j / (M - 1))
    lam = 4 / h[0] ** 2 * math.sin(p * math.pi / (2 * (N - 1))) ** 2 + 4 / h[1] ** 2 * math.sin(
        q * math.pi / (2 * (M - 1))) ** 2
    th = math.acos(1 - 0.5 * dt * dt * lam)
    kc = np.full((N, M), dt * dt)
    n = 150
    _, _, uc = c_oracle.forward(kc, 2, h, n, u_prev=math.cos(th) * mode, u_cur=mode)
    assert np.max(np.abs(uc - math.cos(n * th) * mode)) < 1e-11
    _, _, us = slice_ref.forward(kc, 2, h, n, u_prev=math.cos(th) * mode, u_cur=mode)
    assert np.max(np.abs(us - math.cos(n * th) * mode)) < 1e-11


def test_1d_magic_time_step_is_dalembert():
    """c dt = dx, order 2: u^n_i = (f_{i-n} + f_{i+n}) / 2 exactly for a pulse away from the ends."""
    n = 400
    f = np.zeros(n)
    f[190:211] = np.arange(21) % 7 - 3.0  # small integers: exact in binary
    u_prev = 0.5 * (np.roll(f, 1) + np.roll(f, -1))
    kc = np.ones(n)
    steps = 150
    _, _, uc = c_oracle.forward(kc, 2, (1.0,), steps, u_prev=u_prev, u_cur=f)
    expect = 0.5 * (np.roll(f, steps) + np.roll(f, -steps))
    assert np.array_equal(uc, expect)


def _energy(u_prev, u_cur, kc, order, h):
    d = u_cur - u_prev
    return float(np.sum(d * d / kc) - np.sum(slice_ref.laplacian(u_cur, order, h) * u_prev))


@pytest.mark.parametrize("order", [2, 8])
def test_discrete_energy_conserved_and_sponge_dissipates(order):
    case = make_case((40, 50), order, seed=3, sponge=False, nsrc=0, nrec=0, nsteps=1)
    rng = np.random.default_rng(1)
    inner = (slice(5, 35), slice(5, 45))
    u0 = np.zeros(case["shape"])
    u0[inner] = rng.standard_normal((30, 40))
    up, uc = u0 * 0.9, u0
    kc = case["kc"]
    e0 = _energy(up, uc, kc, order, case["h"])
    _, up2, uc2 = c_oracle.forward(kc, order, case["h"], 400, u_prev=up, u_cur=uc)
    assert abs(_energy(up2, uc2, kc, order, case["h"]) - e0) < 1e-10 * abs(e0)
    prof = [slice_ref.sponge_profile(n, 8, 0.2) for n in case["shape"]]
    es = [e0]
    a, b = up, uc
    for _ in range(5):
        _, a, b = c_oracle.forward(kc, order, case["h"], 80, beta_profiles=prof, u_prev=a, u_cur=b)
        es.append(_energy(a, b, kc, order, case["h"]))
    assert all(y < x for x, y in zip(es, es[1:]))


def test_reciprocity_and_linearity():
    case = make_case((37, 41), 6, seed=7, nsteps=120)
    s, r = (9, 12), (25, 30)
    sig = case["sig"][:, :1]
    prof = case["prof"]
    t1, _, _ = c_oracle.forward(case["kc"], 6, case["h"], 120, beta_profiles=prof, src_nodes=[s], src_sig=sig,
                                rec_nodes=[r])
    t2, _, _ = c_oracle.forward(case["kc"], 6, case["h"], 120, beta_profiles=prof, src_nodes=[r], src_sig=sig,
                                rec_nodes=[s])
    assert rel_l2(t1, t2) < 1e-12
    sig2 = np.random.default_rng(1).standard_normal(sig.shape)
    ta, _, _ = c_oracle.forward(case["kc"], 6, case["h"], 120, src_nodes=[s], src_sig=2 * sig - 3 * sig2, rec_nodes=[r])
    tb, _, _ = c_oracle.forward(case["kc"], 6, case["h"], 120, src_nodes=[s], src_sig=sig, rec_nodes=[r])
    tc, _, _ = c_oracle.forward(case["kc"], 6, case["h"], 120, src_nodes=[s], src_sig=sig2, rec_nodes=[r])
    assert rel_l2(ta, 2 * tb - 3 * tc) < 1e-12


@pytest.mark.parametrize("order", [2, 8])
def test_c_oracle_matches_slice_restatement(order):
    case = make_case((43, 38), order, seed=2, nsteps=150)
    beta64 = slice_ref.beta_field(case["prof"], np.float64)
    t_s, _, u_s = slice_ref.forward(case["kc"], order, case["h"], 150, beta64, case["src"], case["sig"], case["rec"])
    t_c, _, u_c = c_oracle.forward(case["kc"], order, case["h"], 150, beta_profiles=case["prof"],
                                   src_nodes=case["src"], src_sig=case["sig"], rec_nodes=case["rec"])
    assert rel_l2(u_c, u_s) < 1e-13 and rel_l2(t_c, t_s) < 1e-13
    kc32 = case["kc"].astype(np.float32)
    t_s, _, u_s = slice_ref.forward(kc32, order, case["h"], 150, slice_ref.beta_field(case["prof"], np.float32),
                                    case["src"], case["sig"].astype(np.float32), case["rec"])
    t_c, _, u_c = c_oracle.forward(kc32, order, case["h"], 150, beta_profiles=case["prof"], src_nodes=case["src"],
                                   src_sig=case["sig"].astype(np.float32), rec_nodes=case["rec"])
    assert rel_l2(u_c, u_s) < 1e-5 and rel_l2(t_c, t_s) < 1e-5


def test_3d_oracles_agree():
    case = make_case((15, 17, 19), 8, seed=4, nsteps=30)
    beta64 = slice_ref.beta_field(case["prof"], np.float64)
    t_s, _, u_s = slice_ref.forward(case["kc"], 8, case["h"], 30, beta64, case["src"], case["sig"], case["rec"])
    t_c, _, u_c = c_oracle.forward(case["kc"], 8, case["h"], 30, beta_profiles=case["prof"], src_nodes=case["src"],
                                   src_sig=case["sig"], rec_nodes=case["rec"])
    assert rel_l2(u_c, u_s) < 1e-13 and rel_l2(t_c, t_s) < 1e-13


def test_reference_step_equals_general_step_order2():
    """fd_step of 2D_wave_cupy.py:54-62 and the general restatement give the same fp64 field."""
    u0 = slice_ref.gaussian_ic(60, 120)
    dx, dy = 1 / 59, 2 / 119
    dt = slice_ref.reference_dt(dx, dy, 1.0, 0.95)
    a0, a1 = slice_ref.run_reference(u0.copy(), u0.copy(), 100, dt, dx, dy, 1.0)
    kc = np.full(u0.shape, dt * dt)
    up, uc = u0.copy(), u0.copy()
    for bnd in (up, uc):
        bnd[0, :] = bnd[-1, :] = bnd[:, 0] = bnd[:, -1] = 0  # the reference never writes the ring
    a0b, a1b = a0.copy(), a1.copy()
    for arr in (a0b, a1b):
        arr[0, :] = arr[-1, :] = arr[:, 0] = arr[:, -1] = 0
    _, p, c = slice_ref.forward(kc, 2, (dx, dy), 100, u_prev=up, u_cur=uc)
    assert rel_l2(c, a1b) < 1e-13 and rel_l2(p, a0b) < 1e-13


def test_adjoint_gradient_matches_torch_autograd():
    import torch

    case = make_case((21, 26), 6, seed=12, nsteps=45, nrec=4)
    obs = np.random.default_rng(4).standard_normal((45, 4)) * 1e-2
    beta = slice_ref.beta_field(case["prof"], np.float64)
    _, g_kc = slice_ref.adjoint_gradient(case["kc"], 6, case["h"], 45, obs, beta, case["src"], case["sig"], case["rec"])
    g_v = g_kc * 2 * case["v"] * case["dt"] ** 2
    _, gt = torch_ref.misfit_grad_v(torch.as_tensor(case["v"]), case["dt"], 6, case["h"], 45, torch.as_tensor(obs),
                                    beta=torch.as_tensor(beta), src_nodes=case["src"], src_sig=case["sig"],
                                    rec_nodes=case["rec"])
    assert rel_l2(g_v, gt.numpy()) < 1e-12


def test_c_adjoint_matches_slice_adjoint_and_central_fd():
    case = make_case((19, 23), 4, seed=8, nsteps=40, nrec=5)
    obs = np.random.default_rng(9).standard_normal((40, 5)) * 1e-2
    beta = slice_ref.beta_field(case["prof"], np.float64)
    J, g_kc = slice_ref.adjoint_gradient(case["kc"], 4, case["h"], 40, obs, beta, case["src"], case["sig"], case["rec"])
    tr, hist, _, _ = c_oracle.history(case["kc"], 4, case["h"], 40, beta_profiles=case["prof"],
                                      src_nodes=case["src"], src_sig=case["sig"], rec_nodes=case["rec"])
    G, mu = c_oracle.adjoint(case["kc"], 4, case["h"], 40, -(tr - obs), hist, seg=7, beta_profiles=case["prof"],
                             rec_nodes=case["rec"], src_nodes=case["src"])
    for s, node in enumerate(case["src"]):
        G[node] += float(np.dot(mu[:, s], case["sig"][:, s]))
    inner = tuple(slice(1, n - 1) for n in case["shape"])
    gc = np.zeros_like(G)
    gc[inner] = -G[inner] / case["kc"][inner]
    assert rel_l2(gc, g_kc) < 1e-12
    for cell in [(5, 7), (10, 12), (15, 3)]:
        eps = 1e-4 * case["kc"][cell]
        kp, km = case["kc"].copy(), case["kc"].copy()
        kp[cell] += eps
        km[cell] -= eps
        Jp = slice_ref.misfit(slice_ref.forward(kp, 4, case["h"], 40, beta, case["src"], case["sig"], case["rec"])[0], obs)
        Jm = slice_ref.misfit(slice_ref.forward(km, 4, case["h"], 40, beta, case["src"], case["sig"], case["rec"])[0], obs)
        assert (Jp - Jm) / (2 * eps) == pytest.approx(g_kc[cell], rel=1e-6, abs=1e-9 * np.abs(g_kc).max())


def test_ricker_closed_form():
    f0 = 10.0
    t0 = 1 / f0
    assert slice_ref.ricker(np.array([t0]), f0)[0] == 1.0
    tz = t0 + 1 / (math.pi * f0 * math.sqrt(2))
    assert abs(slice_ref.ricker(np.array([tz]), f0)[0]) < 1e-15
    t = np.linspace(t0 - 5 / f0, t0 + 5 / f0, 20001)
    assert abs(np.trapezoid(slice_ref.ricker(t, f0), t)) < 1e-6


def test_sponge_profile_shape():
    p = slice_ref.sponge_profile(100, 20, 0.3)
    assert p[0] == pytest.approx(0.3) and p[-1] == pytest.approx(0.3)
    assert np.all(p[20:80] == 0)
    assert np.all(np.diff(p[:21]) < 0)
    q = slice_ref.sponge_profile(100, 20, 0.3, lo=False)
    assert q[0] == 0 and q[-1] == pytest.approx(0.3)


# ---------------------------------------------------------------- planted mistakes (CPU only)
def test_pins_catch_a_wrong_weight(monkeypatch):
    good = slice_ref.fd_weights

    def bad(order):
        a = list(good(order))
        a[2] = -a[2]  # a plausible sign slip
        return a

    _symbol_check(8)  # the pin passes with the right weights
    monkeypatch.setattr(slice_ref, "fd_weights", bad)
    with pytest.raises(AssertionError):
        _symbol_check(8)


def _symbol_check(order):
    shape, h, xi = (40, 47), (0.7, 1.3), (0.9, 0.4)
    x = np.meshgrid(*[np.arange(n) * hd for n, hd in zip(shape, h)], indexing="ij")
    u = np.cos(xi[0] * x[0] + xi[1] * x[1] + 0.3)
    a = [float(v) for v in TABLE[order]]
    lam = sum(-(a[0] + 2 * sum(a[k] * math.cos(k * q * hd) for k in range(1, len(a)))) / hd ** 2
              for q, hd in zip(xi, h))
    K = order // 2
    inner = tuple(slice(K + 1, n - K - 1) for n in shape)
    lap = slice_ref.laplacian(u, order, h)
    assert np.max(np.abs(lap[inner] + lam * u[inner])) < 1e-11 * lam


def test_pins_catch_transposed_spacing(monkeypatch):
    good = slice_ref.laplacian

    def swapped(u, order, h):
        return good(u, order, tuple(reversed(h)))

    monkeypatch.setattr(slice_ref, "laplacian", swapped)
    with pytest.raises(AssertionError):
        _symbol_check(4)


def test_pins_catch_a_dropped_sponge_term(monkeypatch):
    good_step = slice_ref.step

    def no_prev_damping(u_prev, u_cur, kc, order, h, beta=None, src_nodes=(), src_vals=()):
        out = good_step(u_prev, u_cur, kc, order, h, None, src_nodes, src_vals)
        if beta is not None:
            out = out / (1 + beta)  # drops the (1 - beta) factor on u_prev
        return out

    case = make_case((30, 33), 4, seed=1, nsteps=60)
    beta = slice_ref.beta_field(case["prof"], np.float64)
    t_ok, _, _ = slice_ref.forward(case["kc"], 4, case["h"], 60, beta, case["src"], case["sig"], case["rec"])
    monkeypatch.setattr(slice_ref, "step", no_prev_damping)
    t_bad, _, _ = slice_ref.forward(case["kc"], 4, case["h"], 60, beta, case["src"], case["sig"], case["rec"])
    t_c, _, _ = c_oracle.forward(case["kc"], 4, case["h"], 60, beta_profiles=case["prof"], src_nodes=case["src"],
                                 src_sig=case["sig"], rec_nodes=case["rec"])
    assert rel_l2(t_ok, t_c) < 1e-12
    assert rel_l2(t_bad, t_c) > 1e-3


def test_golden_ricker256_order2_fixture():
    """The committed fixture (tests/golden/make_fixtures.py): the slice-form restatement reproduces
    it bit for bit, and the C restatement (another implementation of the same update) agrees to
    float64 rounding."""
    from pathlib import Path

    from tests.golden.make_fixtures import compute
    from tests.wave_cases import oracle_forward, ricker_256_order2_setup

    gold = np.load(Path(__file__).parent / "golden" / "ricker256_order2.npz")
    now = compute()
    for k in ("traces_every_10", "row_sums", "col_sums", "field_l2"):
        assert np.array_equal(now[k], gold[k]), k
    tr, _, uc = oracle_forward(ricker_256_order2_setup(), np.float64)
    assert rel_l2(tr[::10], gold["traces_every_10"]) < 1e-12
    assert rel_l2(uc.sum(1), gold["row_sums"]) < 1e-12 and rel_l2(uc.sum(0), gold["col_sums"]) < 1e-12
    assert abs(np.linalg.norm(uc) / float(gold["field_l2"]) - 1) < 1e-12


def _round_f32(v: F) -> np.float32:
    """The float32 nearest to the exact rational ``v`` (ties to even)."""
    x = np.float32(float(v))
    cands = [np.nextafter(x, np.float32(-np.inf)), x, np.nextafter(x, np.float32(np.inf))]
    return min(cands, key=lambda c: (abs(F(float(c)) - v), int(np.asarray(c).view(np.uint32)) & 1))


def _fma_exact(a, b, c, dtype):
    v = F(float(a)) * F(float(b)) + F(float(c))
    return _round_f32(v) if dtype == np.float32 else np.float64(float(v))  # int/int division rounds once


def _imaging_exact(mu, lapu, seg, groups, dtype):
    """G of the multi-shot imaging sum in exact arithmetic, one correctly rounded fma at a time:
    per step m (descending) and shot group (a list of shot indices, in the order their terms are
    added), p = fma(mu_s^m, L u_s^{m-1}, p) in ``dtype``; every (m-1) % seg == 0 the partials are
    added into the float64 G in group order and cleared."""
    nshots, nsteps = mu.shape[:2]
    shape = mu.shape[2:]
    G = np.zeros(shape, np.float64)
    P = [np.zeros(shape, dtype) for _ in groups]
    inner = list(np.ndindex(*[n - 2 for n in shape]))
    for m in range(nsteps, 0, -1):
        for q, grp in enumerate(groups):
            for s in grp:
                for idx in inner:
                    x = tuple(i + 1 for i in idx)
                    P[q][x] = _fma_exact(mu[s, m - 1][x], lapu[s, m - 1][x], P[q][x], dtype)
        if (m - 1) % seg == 0:
            for q in range(len(groups)):
                G += P[q].astype(np.float64)
                P[q][...] = 0
    return G


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("group", [1, 2, 4])
def test_multishot_imaging_groups_in_exact_arithmetic(dtype, group):
    """The oracle's multi-shot adjoint: each shot's adjoint is the single-shot one, and the imaging
    partials (one per group of consecutive shots, terms added in shot order, flushed in group
    order) equal an exact-arithmetic accumulation of the same terms; the reversed shot order
    inside a group gives other bits, so the check sees the order."""
    case = make_case((10, 12), 4, seed=41, nsteps=9, nsrc=1, nrec=3)
    S, N = 5, case["nsteps"]
    rng = np.random.default_rng(17)
    kc = case["kc"].astype(dtype)
    srcs = [[(int(rng.integers(1, 9)), int(rng.integers(1, 11)))] for _ in range(S)]
    sigs = rng.standard_normal((S, N, 1)).astype(dtype)
    hist = np.stack([c_oracle.history(kc, 4, case["h"], N, beta_profiles=case["prof"], src_nodes=srcs[s],
                                      src_sig=sigs[s], rec_nodes=case["rec"])[1] for s in range(S)])
    adj = rng.standard_normal((S, N, 3)).astype(dtype)
    G, mu_src, mu = c_oracle.adjoint_shots(kc, 4, case["h"], N, adj, hist, seg=4, group=group,
                                           beta_profiles=case["prof"], rec_nodes=case["rec"], src_nodes=srcs,
                                           keep_mu=True)
    for s in range(S):
        G1, m1 = c_oracle.adjoint(kc, 4, case["h"], N, adj[s], hist[s], seg=4, beta_profiles=case["prof"],
                                  rec_nodes=case["rec"], src_nodes=srcs[s])
        assert np.array_equal(mu_src[s], m1)
    lapu = np.stack([np.stack([c_oracle.lap(hist[s, n], 4, case["h"]) for n in range(N)]) for s in range(S)])
    groups = [list(range(q, min(q + group, S))) for q in range(0, S, group)]
    assert np.array_equal(G, _imaging_exact(mu, lapu, 4, groups, dtype))
    G_one, _ = c_oracle.adjoint_shots(kc, 4, case["h"], N, adj, hist, seg=4, group=1, beta_profiles=case["prof"],
                                      rec_nodes=case["rec"], src_nodes=srcs)
    assert rel_l2(G, G_one) < (1e-6 if dtype == np.float32 else 1e-14)
    if group > 1:
        assert not np.array_equal(G, _imaging_exact(mu, lapu, 4, [g[::-1] for g in groups], dtype))


@pytest.mark.parametrize("order", [2, 8])
def test_oracle_laplacian_export_matches_slice_restatement(order):
    rng = np.random.default_rng(3)
    u = rng.standard_normal((14, 17))
    got = c_oracle.lap(u, order, (0.8, 0.9))
    want = slice_ref.laplacian(u, order, (0.8, 0.9))
    inner = (slice(1, -1), slice(1, -1))
    assert np.allclose(got[inner], want[inner], rtol=1e-12, atol=1e-12)
    assert not got[0].any() and not got[-1].any() and not got[:, 0].any() and not got[:, -1].any()
