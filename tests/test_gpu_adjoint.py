"""GPU adjoint and gradient: bit-exact imaging sum against the C oracle, independence from the
checkpoint spacing, the dot-product test, central finite differences and torch autograd through
the CPU restatement."""

import numpy as np
import pytest

from oracle import c_oracle, torch_ref
from tests.wave_cases import gpu_objects, make_case, rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _plan_grad(case, tdt, seg=0, obs=None, nshots=1):
    from neuralmech_amd import solver

    g, med, src, rec = gpu_objects(case, tdt, nshots=nshots)
    plan = solver.GradientPlan(g, case["nsteps"], seg=seg)
    tr = plan.forward(med, src, case["sig"], rec)
    if obs is None:
        obs = torch.zeros_like(tr)
    resid = tr - torch.as_tensor(obs, device=tr.device, dtype=tr.dtype)
    grad = plan.backward(med, src, rec, resid)
    torch.cuda.synchronize()
    return plan, tr, resid, grad


@pytest.mark.parametrize("dtype", ["f32", "f64"])
@pytest.mark.parametrize("shape,order", [((31, 43), 4), ((40, 36), 8), ((14, 17, 23), 8), ((9, 70, 75), 4),
                                         ((60,), 2)])
def test_imaging_sum_bitexact_vs_oracle(shape, order, dtype):
    print(f"\n[adjoint] shape={shape} order={order} {dtype}", flush=True)
    npdt, tdt = (np.float32, torch.float32) if dtype == "f32" else (np.float64, torch.float64)
    case = make_case(shape, order, seed=21, nsteps=70, nrec=5)
    obs = np.random.default_rng(2).standard_normal((1, 70, 5)) * 1e-3
    plan, tr, resid, grad = _plan_grad(case, tdt, seg=9, obs=obs)
    kc = case["kc"].astype(npdt)
    tr_o, hist, _, _ = c_oracle.history(kc, order, case["h"], 70, beta_profiles=case["prof"],
                                        src_nodes=case["src"], src_sig=case["sig"].astype(npdt),
                                        rec_nodes=case["rec"])
    assert np.array_equal(tr[0].cpu().numpy(), tr_o)
    r = resid[0].cpu().numpy()
    G, mu_src = c_oracle.adjoint(kc, order, case["h"], 70, -r, hist, seg=32, beta_profiles=case["prof"],
                                 rec_nodes=case["rec"], src_nodes=case["src"])
    assert np.array_equal(plan.last_mu_src[0].cpu().numpy(), mu_src)
    assert np.array_equal(plan.last_G.cpu().numpy(), G)
    for s, node in enumerate(case["src"]):
        G[node] += float(np.dot(mu_src[:, s].astype(np.float64), case["sig"][:, s].astype(npdt).astype(np.float64)))
    inner = tuple(slice(1, n - 1) for n in shape)
    ref = np.zeros_like(G)
    ref[inner] = -G[inner] / kc[inner].astype(np.float64)
    assert rel_l2(grad.cpu().numpy(), ref) < 1e-13


def test_gradient_independent_of_checkpoint_spacing():
    case = make_case((47, 39), 8, seed=4, nsteps=61)
    obs = np.zeros((1, 61, len(case["rec"])))
    grads = [_plan_grad(case, torch.float32, seg=s, obs=obs)[3] for s in (1, 2, 7, 13, 61, 0)]
    for gr in grads[1:]:
        assert torch.equal(gr, grads[0])


def test_fastest_spacing_keeps_every_level_when_it_fits():
    """seg = SEG_FASTEST: every level kept (seg 1, no recompute) when the store fits the budget,
    the memory-optimal spacing when it does not; the same gradient bits either way."""
    from neuralmech_amd import solver
    from neuralmech_amd._lib import WaveError

    case = make_case((47, 39), 8, seed=4, nsteps=61)
    obs = np.zeros((1, 61, len(case["rec"])))
    ref = _plan_grad(case, torch.float32, seg=0, obs=obs)
    plan, _, _, grad = _plan_grad(case, torch.float32, seg=solver.SEG_FASTEST, obs=obs)
    assert plan.seg == 1 and plan.nckpt == 61
    assert torch.equal(grad, ref[3])
    g, _, _, _ = gpu_objects(case, torch.float32)
    full = solver.GradientPlan(g, 61, seg=1).bytes
    small = solver.GradientPlan(g, 61, seg=solver.SEG_FASTEST, max_bytes=full - 1)
    assert small.seg == ref[0].seg and small.bytes == ref[0].bytes
    assert solver.GradientPlan(g, 61, seg=solver.SEG_FASTEST, max_bytes=full).seg == 1
    with pytest.raises(WaveError):
        solver.GradientPlan(g, 61, seg=-2)


def test_adjoint_dot_product_fp64():
    """<A s, y> == <s, A^T y> for the signal-to-trace map A, A^T y = -mu at the sources."""
    case = make_case((41, 52), 6, seed=9, nsteps=80, nsrc=3, nrec=7)
    rng = np.random.default_rng(3)
    y = rng.standard_normal((1, 80, 7))
    from neuralmech_amd import solver

    g, med, src, rec = gpu_objects(case, torch.float64)
    plan = solver.GradientPlan(g, 80, seg=10)
    tr = plan.forward(med, src, case["sig"], rec)
    plan.backward(med, src, rec, torch.as_tensor(y))
    lhs = float((tr.cpu().double() * torch.as_tensor(y)).sum())
    ATy = -plan.last_mu_src[0].cpu().numpy()
    rhs = float(np.sum(case["sig"] * ATy))
    assert abs(lhs - rhs) <= 1e-11 * max(abs(lhs), 1.0), (lhs, rhs)


def _misfit_v(case, v, obs, tdt=torch.float64):
    from neuralmech_amd import solver

    c = dict(case, kc=(v * case["dt"]) ** 2)
    g, med, src, rec = gpu_objects(c, tdt)
    tr, _ = solver.run(g, med, case["nsteps"], src=src, sig=case["sig"], rec=rec)
    r = tr[0].cpu().double().numpy() - obs
    return 0.5 * float(np.sum(r * r))


def test_gradient_vs_central_finite_differences_fp64():
    case = make_case((36, 44), 8, seed=13, nsteps=90, nrec=8)
    rng = np.random.default_rng(33)
    obs = rng.standard_normal((90, 8)) * 1e-2
    _, _, _, grad_kc = _plan_grad(case, torch.float64, obs=obs[None])
    v = case["v"]
    grad_v = grad_kc.cpu().numpy() * 2 * v * case["dt"] ** 2
    cells = [(int(rng.integers(2, 34)), int(rng.integers(2, 42))) for _ in range(10)]
    for cell in cells:
        hstep = 1e-4 * v[cell]
        vp = v.copy()
        vp[cell] += hstep
        vm = v.copy()
        vm[cell] -= hstep
        fd = (_misfit_v(case, vp, obs) - _misfit_v(case, vm, obs)) / (2 * hstep)
        assert abs(fd - grad_v[cell]) <= 1e-6 * np.max(np.abs(grad_v)) + 1e-7 * abs(fd), (cell, fd, grad_v[cell])


def test_gradient_vs_torch_autograd_on_cpu_restatement():
    case = make_case((24, 29), 4, seed=17, nsteps=50, nrec=5)
    obs = np.random.default_rng(5).standard_normal((50, 5)) * 1e-2
    _, _, _, grad_kc = _plan_grad(case, torch.float64, obs=obs[None])
    grad_v = grad_kc.cpu().numpy() * 2 * case["v"] * case["dt"] ** 2
    from oracle import slice_ref

    beta = torch.as_tensor(slice_ref.beta_field(case["prof"], np.float64))
    _, gt = torch_ref.misfit_grad_v(torch.as_tensor(case["v"]), case["dt"], 4, case["h"], 50, torch.as_tensor(obs),
                                    beta=beta, src_nodes=case["src"], src_sig=case["sig"], rec_nodes=case["rec"])
    assert rel_l2(grad_v, gt.numpy()) < 1e-10


def test_batched_gradient_is_sum_of_shot_gradients():
    from neuralmech_amd import solver

    case = make_case((33, 51), 8, seed=23, nsteps=40, nrec=6)
    shots = [[(5, 10)], [(5, 25)], [(5, 40)]]
    sig = np.random.default_rng(6).standard_normal((3, 40, 1))
    g, med, _, _ = gpu_objects(case, torch.float64, nshots=3)
    src = solver.PointSet(g, shots, per_shot=True)
    rec = solver.PointSet(g, case["rec"])
    plan = solver.GradientPlan(g, 40)
    tr = plan.forward(med, src, torch.as_tensor(sig), rec)
    total = plan.backward(med, src, rec, tr)
    acc = torch.zeros_like(total)
    for s in range(3):
        c1 = dict(case, src=shots[s], sig=sig[s])
        acc += _plan_grad(c1, torch.float64, obs=np.zeros((1, 40, 6)))[3]
    assert rel_l2(total.cpu().numpy(), acc.cpu().numpy()) < 1e-13


@pytest.mark.parametrize("variant,rows", [(0, 2), (1, 1), (1, 8), (2, 2), (3, 8)])
def test_adjoint_launch_shapes_bitexact(variant, rows):
    case = make_case((45, 300), 8, seed=29, nsteps=30, nrec=5)
    from neuralmech_amd import solver

    base = _plan_grad(case, torch.float32, seg=6)
    g, med, src, rec = gpu_objects(case, torch.float32)
    g.set_launch(variant, rows)
    try:
        plan = solver.GradientPlan(g, 30, seg=6)
        tr = plan.forward(med, src, case["sig"], rec)
        grad = plan.backward(med, src, rec, tr)
        torch.cuda.synchronize()
    finally:
        g.set_launch(0, 0)
    assert torch.equal(tr, base[1]) and torch.equal(grad, base[3])


@pytest.mark.parametrize("variant,rows", [(0, 4), (2, 4), (2, 16), (2, 128)])
def test_adjoint_3d_launch_shapes_bitexact(variant, rows):
    """3D gradient plan (forward, recompute, adjoint with imaging) under forced launch shapes,
    streaming access included, against the default shape: same traces and gradient bits."""
    from neuralmech_amd import solver

    case = make_case((12, 40, 75), 8, seed=31, nsteps=24, nrec=5)
    base = _plan_grad(case, torch.float32, seg=5)
    g, med, src, rec = gpu_objects(case, torch.float32)
    g.set_launch(variant, rows)
    try:
        plan = solver.GradientPlan(g, 24, seg=5)
        tr = plan.forward(med, src, case["sig"], rec)
        grad = plan.backward(med, src, rec, tr)
        torch.cuda.synchronize()
    finally:
        g.set_launch(0, 0)
    assert torch.equal(tr, base[1]) and torch.equal(grad, base[3])



def test_two_outstanding_forwards_keep_their_own_plans():
    """wave_traces keeps its gradient plan for the next call; two forwards alive at once (one
    backward through both) must not share it: gradients equal one-at-a-time runs bit for bit, and
    a later call reuses the kept plan with the same result."""
    from neuralmech_amd import autograd

    case = make_case((40, 60), 4, seed=5, nsteps=30, nrec=4)
    g, med, src, rec = gpu_objects(case, torch.float64)
    sig = torch.as_tensor(case["sig"], device="cuda")
    v0 = torch.as_tensor(case["v"], dtype=torch.float64, device="cuda")

    def grad_of(v):
        w = v.clone().requires_grad_(True)
        (autograd.wave_traces(w, g, src, sig, rec, case["prof"]) ** 2).sum().backward()
        return w.grad

    autograd.clear_plans()
    a = v0.clone().requires_grad_(True)
    b = (1.1 * v0).requires_grad_(True)
    da = autograd.wave_traces(a, g, src, sig, rec, case["prof"])
    db = autograd.wave_traces(b, g, src, sig, rec, case["prof"])
    ((da ** 2).sum() + (db ** 2).sum()).backward()
    ga, gb = grad_of(v0), grad_of(1.1 * v0)
    assert torch.equal(a.grad, ga) and torch.equal(b.grad, gb)
    assert torch.equal(grad_of(v0), ga)  # the kept plan again
    autograd.clear_plans()


@pytest.mark.parametrize("variant,rows", [(4, 2), (5, 8), (6, 4), (7, 1)])
def test_shot_fastest_tile_order_bitexact(variant, rows):
    """Several shots on one grid: the traces and the velocity gradient with the tiles taken
    shot-fastest (WFD_VARIANT_SHOTS = 4) equal those of the same halo / access variant in the
    default order bit for bit (only the order in which tiles run changes), for 1..8 rows."""
    from neuralmech_amd import solver

    print(f"\n[shot order] variant={variant} rows={rows}", flush=True)
    case = make_case((41, 530), 8, seed=31, nsteps=36, nrec=7)
    shots = [[(5, 10 + 120 * s)] for s in range(4)]
    sig = np.random.default_rng(8).standard_normal((4, 36, 1))
    out = {}
    for v in (variant & 3, variant):
        g, med, _, _ = gpu_objects(case, torch.float32, nshots=4)
        src = solver.PointSet(g, shots, per_shot=True)
        rec = solver.PointSet(g, case["rec"])
        g.set_launch(v, rows)
        try:
            assert g.launch() == (v, rows)
            plan = solver.GradientPlan(g, 36, seg=6)
            tr = plan.forward(med, src, torch.as_tensor(sig), rec)
            grad = plan.backward(med, src, rec, tr)
            torch.cuda.synchronize()
        finally:
            g.set_launch(0, 0)
        out[v] = (tr, grad)
    (tr0, g0), (tr1, g1) = out[variant & 3], out[variant]
    assert torch.isfinite(g0).all() and g0.abs().max() > 0
    assert torch.equal(tr0, tr1) and torch.equal(g0, g1)


@pytest.mark.parametrize("dtype", ["f32", "f64"])
@pytest.mark.parametrize("shape,order,nshots,group", [((33, 51), 8, 6, 4), ((40, 36), 4, 3, 4),
                                                      ((29, 130), 8, 4, 4), ((22, 300), 6, 9, 4),
                                                      ((12, 20, 23), 8, 3, 1), ((60,), 2, 3, 1)])
def test_multishot_imaging_sum_bitexact_vs_oracle(shape, order, nshots, group, dtype):
    """Gradient plan over several shots in one grid: every shot's traces and adjoint source
    records, and the imaging sum of all shots, bit-exact against the oracle's multi-shot adjoint.
    2D grids keep one partial per group of 4 shots (terms added in shot order, the last group
    with fewer live shots when 4 does not divide the count, under every launch shape); 1D and 3D
    one per shot."""
    from neuralmech_amd import solver

    print(f"\n[adjoint shots] shape={shape} order={order} shots={nshots} {dtype}", flush=True)
    npdt, tdt = (np.float32, torch.float32) if dtype == "f32" else (np.float64, torch.float64)
    N = 37
    case = make_case(shape, order, seed=43, nsteps=N, nsrc=1, nrec=4)
    rng = np.random.default_rng(44)
    srcs = [[tuple(int(rng.integers(1, n - 1)) for n in shape)] for _ in range(nshots)]
    sig = rng.standard_normal((nshots, N, 1))
    obs = rng.standard_normal((nshots, N, 4)) * 1e-3
    g, med, _, _ = gpu_objects(case, tdt, nshots=nshots)
    src = solver.PointSet(g, srcs, per_shot=True)
    rec = solver.PointSet(g, case["rec"])
    plan = solver.GradientPlan(g, N, seg=5)
    tr = plan.forward(med, src, torch.as_tensor(sig), rec)
    resid = tr - torch.as_tensor(obs, device=tr.device, dtype=tr.dtype)
    plan.backward(med, src, rec, resid)
    torch.cuda.synchronize()
    kc = case["kc"].astype(npdt)
    hist = []
    for s in range(nshots):
        tr_o, h_s, _, _ = c_oracle.history(kc, order, case["h"], N, beta_profiles=case["prof"], src_nodes=srcs[s],
                                           src_sig=sig[s].astype(npdt), rec_nodes=case["rec"])
        assert np.array_equal(tr[s].cpu().numpy(), tr_o)
        hist.append(h_s)
    G, mu_src = c_oracle.adjoint_shots(kc, order, case["h"], N, -resid.cpu().numpy(), np.stack(hist), seg=32,
                                       group=group, beta_profiles=case["prof"], rec_nodes=case["rec"],
                                       src_nodes=srcs)
    assert np.array_equal(plan.last_mu_src.cpu().numpy(), mu_src)
    assert np.array_equal(plan.last_G.cpu().numpy(), G)
    if len(shape) != 2:
        return
    for seg in (solver.SEG_FASTEST, 1, N, 7):  # every level kept, one segment, a ragged last segment
        plan = solver.GradientPlan(g, N, seg=seg)
        plan.forward(med, src, torch.as_tensor(sig), rec)
        plan.backward(med, src, rec, resid)
        torch.cuda.synchronize()
        assert np.array_equal(plan.last_G.cpu().numpy(), G), seg
    shapes = [(0, 2), (3, 8), (6, 4)] if dtype == "f32" else [(0, 4), (3, 4), (6, 4)]  # f64 marches 4 rows
    for v, rows in shapes:  # every 2D launch shape groups the same way
        g.set_launch(v, rows)
        try:
            plan = solver.GradientPlan(g, N, seg=5)
            plan.forward(med, src, torch.as_tensor(sig), rec)
            plan.backward(med, src, rec, resid)
            torch.cuda.synchronize()
        finally:
            g.set_launch(0, 0)
        assert np.array_equal(plan.last_G.cpu().numpy(), G), (v, rows)
