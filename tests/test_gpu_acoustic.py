"""Variable mass / stiffness model on the GPU (wfd_ac_forward / wfd_ac_backward) against the
operation-for-operation restatement in oracle/acoustic_ref.py (bit-exact, fp32 and fp64), the
float64 full-history transpose (gradient within 1e-3 relative L2 from fp32 runs), and the cuwave
facade (response_gradient against central differences of response)."""

import math

import numpy as np
import pytest

from oracle import acoustic_ref as ar
from oracle import slice_ref

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _smooth(shape, seed, lo, hi):
    rng = np.random.default_rng(seed)
    x = np.linspace(0, 1, shape[0])[:, None]
    y = np.linspace(0, 1, shape[1])[None, :]
    f = np.zeros(shape)
    for _ in range(4):
        a, b, c, d = rng.uniform(0.5, 3.0, 4)
        f += np.sin(a * math.pi * x + c) * np.cos(b * math.pi * y + d)
    f = (f - f.min()) / (f.max() - f.min())
    return lo + (hi - lo) * f


def _profile(n, width, beta):
    p = np.zeros(n)
    for e in range(width):
        v = beta * ((width - e) / width) ** 2
        p[1 + e] = max(p[1 + e], v)
        p[n - 2 - e] = max(p[n - 2 - e], v)
    return p


def _case(dtype, closure, shape=(37, 70), N=64, sponge=True, seed=3, order=2):
    from neuralmech_amd import acoustic as ac
    from neuralmech_amd.solver import PointSet, WaveGrid

    h = (0.05, 0.04)
    m = _smooth(shape, seed, 0.6, 2.5)
    s = _smooth(shape, seed + 1, 0.3, 1.8)
    dt = 0.45 * min(h) / float(np.sqrt((s / m).max())) / math.sqrt(2)
    beta = (_profile(shape[0], 6, 0.1), _profile(shape[1], 7, 0.12)) if sponge else None
    grid = WaveGrid(shape, h, dt, order, dtype, nshots=2, device="cuda")
    med = ac.AcMedium(grid, m, s, closure, beta)
    # shot 0: two sources on one node (list order matters) and one near a wall; shot 1: one source
    src_nodes = [[(15, 30), (15, 30), (8, 66)], [(20, 12)]]
    t = np.arange(N) * dt
    sig = np.zeros((2, N, 3))
    for i in range(3):
        sig[0, :, i] = slice_ref.ricker(t - 3 * i * dt, 1.0 / (14 * dt)) * (1.0 + 0.3 * i)
    sig[1, :, 0] = slice_ref.ricker(t, 1.0 / (10 * dt))
    rec_nodes = [(10, 10), (20, 40), (25, 64), (16, 30), (9, 63)]
    strip_nodes = med.strip_nodes()
    src = PointSet(grid, src_nodes, per_shot=True)
    rec = PointSet(grid, rec_nodes)
    strip = PointSet(grid, [tuple(int(c) for c in p) for p in strip_nodes]) if len(strip_nodes) else None
    npdt = np.float32 if dtype == torch.float32 else np.float64
    coeff = [f.cpu().numpy().astype(npdt) for f in med.bonds()]
    prof = (None, None) if med.profiles is None else tuple(med.profiles)
    return dict(ac=ac, grid=grid, med=med, src=src, rec=rec, strip=strip, sig=sig, N=N, npdt=npdt, coeff=coeff,
                prof=prof, src_nodes=src_nodes, rec_nodes=rec_nodes, strip_nodes=strip_nodes, shape=shape)


@pytest.mark.parametrize("order", [2, 4, 8])
@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("closure", ["neumann", "dirichlet"])
def test_forward_bitexact_with_oracle(dtype, closure, order):
    C = _case(getattr(torch, dtype), closure, order=order)
    ac = C["ac"]
    tr, store, state = ac.forward(C["grid"], C["med"], C["N"], C["src"], torch.as_tensor(C["sig"]), C["rec"],
                                  C["strip"])
    torch.cuda.synchronize()
    k, cn, cw = C["coeff"]
    for shot in range(2):
        ns = len(C["src_nodes"][shot])
        otr, ostore, (opp, opc) = ar.forward_np(k, cn, cw, C["N"], C["prof"], C["src_nodes"][shot],
                                                C["sig"][shot, :, :ns], C["rec_nodes"], C["strip_nodes"],
                                                dtype=C["npdt"])
        assert np.array_equal(tr[shot].cpu().numpy(), otr)
        assert np.array_equal(store[shot].cpu().numpy(), ostore)
        assert np.array_equal(C["grid"].nodes(state[1])[shot].cpu().numpy(), opc)
        assert np.array_equal(C["grid"].nodes(state[0])[shot].cpu().numpy(), opp)
    assert float(tr.abs().max()) > 0


@pytest.mark.parametrize("order", [2, 4, 8])
@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("closure", ["neumann", "dirichlet"])
def test_backward_bitexact_with_oracle(dtype, closure, order):
    C = _case(getattr(torch, dtype), closure, N=70, order=order)
    ac = C["ac"]
    sig = torch.as_tensor(C["sig"])
    tr, store, state = ac.forward(C["grid"], C["med"], C["N"], C["src"], sig, C["rec"], C["strip"])
    obs = (0.5 * tr).to(tr.dtype)
    dJ = tr - obs
    G, mu_src = ac.backward(C["grid"], C["med"], C["N"], state, C["src"], sig, C["rec"], dJ, C["strip"], store)
    torch.cuda.synchronize()
    k, cn, cw = C["coeff"]
    for shot in range(2):
        ns = len(C["src_nodes"][shot])
        otr, ostore, last = ar.forward_np(k, cn, cw, C["N"], C["prof"], C["src_nodes"][shot],
                                          C["sig"][shot, :, :ns], C["rec_nodes"], C["strip_nodes"], dtype=C["npdt"])
        adj = np.zeros((C["N"] + 1, len(C["rec_nodes"])), C["npdt"])
        adj[1:] = -dJ[shot].cpu().numpy()
        OG, omu = ar.backward_np(k, cn, cw, C["N"], last, C["prof"], C["src_nodes"][shot], C["sig"][shot, :, :ns],
                                 C["rec_nodes"], adj, C["strip_nodes"], ostore, dtype=C["npdt"])
        assert np.isfinite(OG).all()
        assert np.array_equal(G[:, shot].cpu().numpy(), OG)
        assert np.array_equal(mu_src[shot, :, :ns].cpu().numpy(), omu)
    assert float(G.abs().max()) > 0


@pytest.mark.parametrize("order", [2, 4])
def test_fp32_gradient_against_float64_transpose(order):
    """Gradient with respect to mass and stiffness from an fp32 run (reconstruction adjoint)
    within 1e-3 relative L2 of the float64 full-history transpose."""
    C = _case(torch.float32, "neumann", shape=(61, 90), N=300, order=order)
    ac = C["ac"]
    sig = torch.as_tensor(C["sig"])
    k64, cn64, cw64 = (a.numpy() for a in ac.face_operator(C["med"].mass.cpu(), C["med"].stiff.cpu(),
                                                             C["grid"].spacing, C["grid"].dt, "neumann", order))
    prof = C["prof"]
    ns = len(C["src_nodes"][0])
    obs = [0.8 * ar.traces_exact(k64, cn64, cw64, C["N"], prof, C["src_nodes"][s],
                                 C["sig"][s, :, :len(C["src_nodes"][s])], C["rec_nodes"]) for s in range(2)]
    tr, store, state = ac.forward(C["grid"], C["med"], C["N"], C["src"], sig, C["rec"], C["strip"])
    dJ = tr.double() - torch.as_tensor(np.stack(obs), device="cuda")
    G, mu_src = ac.backward(C["grid"], C["med"], C["N"], state, C["src"], sig, C["rec"], dJ, C["strip"], store)
    dm, ds = ac.material_gradient(C["med"], G, mu_src, sig, C["src"])
    # float64 reference for shot 0 through the same coefficient map
    tr64 = ar.traces_exact(k64, cn64, cw64, C["N"], prof, C["src_nodes"][0], C["sig"][0, :, :ns], C["rec_nodes"])
    ek, ecn, ecw = ar.adjoint_exact(k64, cn64, cw64, C["N"], prof, C["src_nodes"][0], C["sig"][0, :, :ns],
                                    C["rec_nodes"], tr64 - obs[0])
    m = C["med"].mass.cpu().clone().requires_grad_(True)
    s = C["med"].stiff.cpu().clone().requires_grad_(True)
    k, cn, cw = ac.face_operator(m, s, C["grid"].spacing, C["grid"].dt, "neumann", order)
    und = torch.as_tensor(C["med"].undamped)
    ekt = torch.where(und, torch.as_tensor(ek), torch.zeros(()))
    rm, rs = torch.autograd.grad([k, cn, cw], [m, s], [ekt, torch.as_tensor(ecn), torch.as_tensor(ecw)])
    rm = torch.where(und, rm, torch.zeros(()))
    rs = torch.where(und, rs, torch.zeros(()))
    for got, ref in ((dm[0].cpu(), rm), (ds[0].cpu(), rs)):
        err = float(torch.linalg.norm(got - ref) / torch.linalg.norm(ref))
        print(f"fp32 reconstruction gradient vs fp64 transpose: relative L2 {err:.2e}")
        assert err < 1e-3


def test_facade_response_gradient_central_differences():
    """topopt_acoustic2D.py-style call: AcousticWave, sponge, point source, box sensors, energy
    objective, response_gradient with reconstruction_sensitivity (float64) against central
    differences of response in the indicator."""
    from dataclasses import replace

    from neuralmech_amd import boundary, geometry, scalar, sensitivity, signals, utils, wave

    Nx, width, origin, region = boundary.pad_for_sponge((40, 30), (0.4, 0.4), 2.0)
    dx = (0.4, 0.4)
    rho1, rho2, k1, k2 = 1.2, 30.0, 1.4e5, 9.7e5
    cmax = max(math.sqrt(k1 / rho1), math.sqrt(k2 / rho2))
    dt = 0.5 * wave.stable_dt(dx, cmax, 2)
    N = 260
    sim = scalar.AcousticWave(Nx, dx, N, dt, (4, 64), precision="float64", space_order=2, rho1=rho1, rho2=rho2,
                              kappa1=k1, kappa2=k2)
    sim = replace(sim, damping=boundary.sponge(sim, None, width, 0.2))
    t = np.arange(N) * dt
    src = utils.point_source(sim, (origin[0] + 2.0, origin[1] + 6.0), signals.sineburst(t, 1e3, 150.0, 2))
    coords = wave.grid_coords(Nx, dx)
    sensors = geometry.nodes(geometry.box(coords, (origin[0] + 13.0, origin[1] + 6.0), (2.0, 2.0)))
    rng = np.random.default_rng(1)
    gamma = torch.as_tensor(rng.uniform(0.2, 0.8, Nx), dtype=torch.float64)
    objective = utils.energy(sim)
    cost, grad = utils.response_gradient(sim, src, gamma, sensors, objective,
                                         adjoint=sensitivity.reconstruction_sensitivity)
    assert cost > 0 and torch.isfinite(grad).all()
    strip, und = sensitivity.reconstruction_nodes(sim)
    pts = np.argwhere(und)
    picks = pts[rng.choice(len(pts), 5, replace=False)]
    for i, j in picks:
        eps = 1e-4
        gp, gm = gamma.clone(), gamma.clone()
        gp[i, j] += eps
        gm[i, j] -= eps
        fd = (utils.response(sim, src, gp, sensors, objective)[0] - utils.response(sim, src, gm, sensors,
                                                                                     objective)[0]) / (2 * eps)
        g = float(grad[i, j])
        assert abs(fd - g) <= 1e-5 * max(abs(fd), 1e-3 * float(grad.abs().max())), (i, j, fd, g)
    # simulate dispatches to the same forward; traces at the sensors match response's
    u, tr = wave.simulate(sim, src, gamma, sensors)
    c2, tr2 = utils.response(sim, src, gamma, sensors, objective)
    assert torch.equal(tr, tr2) and u.shape == tuple(Nx)


def test_autograd_function_matches_material_gradient():
    """acoustic_traces (torch.autograd.Function over the reconstruction adjoint) in float64 against
    central differences of the traces' misfit in the stiffness at a few nodes."""
    from neuralmech_amd import acoustic as ac
    from neuralmech_amd.solver import PointSet, WaveGrid

    shape, h = (30, 41), (0.05, 0.05)
    m = _smooth(shape, 7, 0.8, 2.0)
    s = _smooth(shape, 8, 0.5, 1.5)
    dt = 0.5 * min(h) / float(np.sqrt((s / m).max())) / math.sqrt(2)
    N = 120
    grid = WaveGrid(shape, h, dt, 2, torch.float64, device="cuda")
    beta = (_profile(shape[0], 5, 0.1), _profile(shape[1], 5, 0.1))
    src = PointSet(grid, [(14, 20)])
    rec = PointSet(grid, [(8, 10), (20, 30)])
    sig = torch.as_tensor(slice_ref.ricker(np.arange(N) * dt, 1.0 / (12 * dt))[None, :, None])
    mt = torch.as_tensor(m, device="cuda").requires_grad_(True)
    st = torch.as_tensor(s, device="cuda").requires_grad_(True)
    tr = ac.acoustic_traces(mt, st, grid, src, sig, rec, "neumann", beta)
    J = (tr ** 2).sum()
    J.backward()
    for (i, j) in [(14, 21), (10, 15), (20, 25)]:
        eps = 1e-6
        sp, sm = s.copy(), s.copy()
        sp[i, j] += eps
        sm[i, j] -= eps
        jp = float((ac.acoustic_traces(m, sp, grid, src, sig, rec, "neumann", beta) ** 2).sum())
        jm = float((ac.acoustic_traces(m, sm, grid, src, sig, rec, "neumann", beta) ** 2).sum())
        fd = (jp - jm) / (2 * eps)
        g = float(st.grad[i, j])
        assert abs(fd - g) <= 1e-5 * max(abs(fd), 1e-3 * float(st.grad.abs().max())), (i, j, fd, g)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_bitexact_on_every_rows_per_wave(dtype):
    """16 shots of 258 x 520 nodes, forward and backward, order 4, with each wave marching 1, 2, 4
    and 8 rows (forced; the automatic choice takes 2 forward and 1 backward here, 4 / 8 on large
    grids); shots 0 and 15 against the restatement every time."""
    from neuralmech_amd import acoustic as ac
    from neuralmech_amd.solver import PointSet, WaveGrid

    shape, order, N, h = (258, 520), 4, 40, (0.05, 0.04)
    tdt = getattr(torch, dtype)
    m = _smooth(shape, 9, 0.6, 2.5)
    s = _smooth(shape, 10, 0.3, 1.8)
    dt = 0.45 * min(h) / float(np.sqrt((s / m).max())) / math.sqrt(2)
    beta = (_profile(shape[0], 10, 0.1), _profile(shape[1], 12, 0.1))
    grid = WaveGrid(shape, h, dt, order, tdt, nshots=16, device="cuda")
    med = ac.AcMedium(grid, m, s, "neumann", beta)
    src_nodes, rec_nodes = [(100, 250), (30, 400)], [(40, 40), (200, 300), (250, 510)]
    t = np.arange(N) * dt
    sig = np.stack([slice_ref.ricker(t, 1.0 / (10 * dt)), slice_ref.ricker(t - 4 * dt, 1.0 / (12 * dt))], 1)
    src, rec = PointSet(grid, src_nodes), PointSet(grid, rec_nodes)
    sn = med.strip_nodes()
    strip = PointSet(grid, [tuple(int(c) for c in p) for p in sn])
    sg = torch.as_tensor(sig)
    npdt = np.float32 if dtype == "float32" else np.float64
    k, cn, cw = (f.cpu().numpy().astype(npdt) for f in med.bonds())
    prof = tuple(med.profiles)
    otr, ostore, last = ar.forward_np(k, cn, cw, N, prof, src_nodes, sig, rec_nodes, sn, dtype=npdt)
    adj = np.zeros((N + 1, len(rec_nodes)), npdt)
    adj[1:] = -0.5 * otr
    OG, omu = ar.backward_np(k, cn, cw, N, last, prof, src_nodes, sig, rec_nodes, adj, sn, ostore, dtype=npdt)
    try:
        for rows in (1, 2, 4, 8, 0):
            print(f"\n[acoustic] 16 shots 258x520 {dtype} rows={rows}", flush=True)
            ac.set_rows(rows)
            tr, store, state = ac.forward(grid, med, N, src, sg, rec, strip)
            dJ = torch.as_tensor(0.5 * otr, device=tr.device).expand_as(tr).contiguous()
            G, mu_src = ac.backward(grid, med, N, state, src, sg, rec, dJ, strip, store)
            torch.cuda.synchronize()
            for shot in (0, 15):
                assert np.array_equal(tr[shot].cpu().numpy(), otr), rows
                assert np.array_equal(G[:, shot].cpu().numpy(), OG), rows
                assert np.array_equal(mu_src[shot].cpu().numpy(), omu), rows
    finally:
        ac.set_rows(0)
