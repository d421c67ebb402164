"""GPU tests of the uneven shot batch (one persistent launch for many shots of different grids,
orders and step counts): every output bit-identical to the single-shot gradient plan and to the
C oracle's forward traces, independent of the schedule, mode, graph replay and tile shapes."""

import numpy as np
import pytest

from oracle import c_oracle, slice_ref

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

# (shape, order, nsteps, seg, sponge)
SPECS = [((40, 56), 2, 37, 0, True), ((64, 64), 4, 80, 0, False), ((48, 100), 6, 65, 7, True),
         ((70, 40), 8, 100, 1, True), ((33, 300), 4, 50, 50, True), ((300, 36), 8, 45, 0, True)]


def _shots(dtype=np.float32, specs=SPECS, seed=0):
    from neuralmech_amd.batch import Shot

    out = []
    for k, (shape, order, N, seg, sponge) in enumerate(specs):
        rng = np.random.default_rng(seed + 17 * k)
        h = (0.9, 1.1)
        v = rng.uniform(1.0, 2.0, shape)
        dt = 0.6 * slice_ref.stable_dt(h, 2.0, order)
        ns, nr = 1 + k % 2, 5 + k
        pts = list(zip(rng.integers(1, shape[0] - 1, ns + nr), rng.integers(1, shape[1] - 1, ns + nr)))
        src = [tuple(int(x) for x in p) for p in pts[:ns]]
        rec = [tuple(int(x) for x in p) for p in pts[ns:]]
        sig = rng.standard_normal((N, ns)).astype(dtype)
        obs = (1e-3 * rng.standard_normal((N, nr))).astype(dtype)
        prof = [slice_ref.sponge_profile(n, max(1, n // 5), 0.15) for n in shape] if sponge else None
        out.append(Shot(shape, h, dt, order, v, src, sig, rec, obs, N, prof, seg))
    return out


def _plan_reference(sh, tdtype):
    """The single-shot gradient plan on the same shot, residual taken in the grid dtype."""
    from neuralmech_amd import solver

    g = solver.WaveGrid(sh.shape, sh.spacing, sh.dt, sh.order, tdtype)
    med = solver.Medium(g, velocity=torch.as_tensor(sh.velocity), beta_profiles=sh.beta_profiles)
    src = solver.PointSet(g, sh.src_nodes)
    rec = solver.PointSet(g, sh.rec_nodes)
    plan = solver.GradientPlan(g, sh.nsteps, seg=sh.seg)
    sig = torch.as_tensor(sh.signal)[None].to("cuda", tdtype).contiguous()
    tr = plan.forward(med, src, sig, rec)
    resid = tr - torch.as_tensor(sh.observed)[None].to("cuda", tdtype)
    grad_kc = plan.backward(med, src, rec, resid)
    v = torch.as_tensor(sh.velocity, dtype=torch.float64, device="cuda")
    return dict(traces=tr, adj=-resid, mu=plan.last_mu_src, G=plan.last_G, grad=grad_kc * 2.0 * v * g.dt ** 2)


def _outputs(b):
    return [(b.traces[k].clone(), b.adj[k].clone(), b.mu_src[k].clone(), b.imaging_sum(k).clone())
            for k in range(len(b.shots))]


def _assert_same(o1, o2):
    for a, b in zip(o1, o2):
        for x, y in zip(a, b):
            assert torch.equal(x, y)


@pytest.fixture(scope="module")
def f32_batch():
    from neuralmech_amd.batch import ShotBatch

    shots = _shots()
    b = ShotBatch(shots, torch.float32)
    b.run(graph=False)
    b.status()
    return shots, b, _outputs(b)


def test_batch_matches_single_shot_plan(f32_batch):
    shots, b, _ = f32_batch
    grads = b.gradients()
    for k, sh in enumerate(shots):
        ref = _plan_reference(sh, torch.float32)
        print("shot", k, sh.shape, "order", sh.order, "steps", sh.nsteps, flush=True)
        assert torch.equal(b.traces[k], ref["traces"])
        assert torch.equal(b.adj[k], ref["adj"])
        S = len(sh.src_nodes)
        assert torch.equal(b.mu_src[k][:, :, :S], ref["mu"])
        assert torch.equal(b.imaging_sum(k), ref["G"])
        assert torch.equal(grads[k][1], ref["grad"])
        r = ref["traces"].double() - torch.as_tensor(sh.observed, device="cuda").double()
        assert grads[k][0] == pytest.approx(0.5 * float((r * r).sum()), rel=1e-12)
        assert float(ref["G"].abs().max()) > 0


def test_batch_traces_match_oracle(f32_batch):
    shots, b, _ = f32_batch
    for k in (0, 3, 5):
        sh = shots[k]
        kc = ((np.asarray(sh.velocity) * sh.dt) ** 2).astype(np.float32)
        tr, _, _ = c_oracle.forward(kc, sh.order, sh.spacing, sh.nsteps, beta_profiles=sh.beta_profiles,
                                    src_nodes=sh.src_nodes, src_sig=sh.signal, rec_nodes=sh.rec_nodes,
                                    dtype=np.float32, shape=sh.shape)
        assert np.array_equal(b.traces[k][0].cpu().numpy(), tr)


def test_batch_modes_graph_and_schedules_agree(f32_batch):
    from neuralmech_amd.batch import ShotBatch

    shots, b, ref = f32_batch
    for kw in [dict(mode="phases"), dict(max_active_cells=1), dict(max_active_cells=1 << 40, blocks_per_cu=1)]:
        print("batch options", kw, flush=True)
        b2 = ShotBatch(shots, torch.float32, **kw)
        b2.run(graph=False)
        b2.status()
        _assert_same(_outputs(b2), ref)
    print("graph replays", flush=True)
    b.run(graph=True)
    b.status()
    _assert_same(_outputs(b), ref)
    b.run(graph=True)  # a second replay of the same graph
    b.status()
    _assert_same(_outputs(b), ref)


def test_batch_tile_shapes_and_tuner_keep_bits(f32_batch):
    from neuralmech_amd.batch import ShotBatch

    shots, _, ref = f32_batch
    b = ShotBatch(shots, torch.float32)
    for variant in (0, 1):
        for rows in (2, 8):
            print("tile shape", variant, rows, flush=True)
            b.set_config(variant, rows)
            b.run(graph=False)
            b.status()
            _assert_same(_outputs(b), ref)
    cfg = b.tune(steps=4)
    assert all(v in (0, 1) and r in (2, 8) for v, r in cfg)
    b.run(graph=True)
    b.status()
    _assert_same(_outputs(b), ref)


def test_batch_float64_gradient_matches_central_differences():
    from neuralmech_amd import solver
    from neuralmech_amd.batch import ShotBatch

    specs = [((30, 44), 4, 60, 0, True), ((52, 28), 8, 45, 5, False)]
    shots = _shots(np.float64, specs, seed=5)
    b = ShotBatch(shots, torch.float64)
    b.run(graph=False)
    b.status()
    grads = b.gradients()
    for k, sh in enumerate(shots):
        ref = _plan_reference(sh, torch.float64)
        assert torch.equal(b.imaging_sum(k), ref["G"]) and torch.equal(grads[k][1], ref["grad"])

        def J(v):
            g = solver.WaveGrid(sh.shape, sh.spacing, sh.dt, sh.order, torch.float64)
            med = solver.Medium(g, velocity=torch.as_tensor(v), beta_profiles=sh.beta_profiles)
            tr, _ = solver.run(g, med, sh.nsteps, src=solver.PointSet(g, sh.src_nodes),
                               sig=torch.as_tensor(sh.signal)[None], rec=solver.PointSet(g, sh.rec_nodes))
            r = tr[0] - torch.as_tensor(sh.observed, device="cuda")
            return 0.5 * float((r * r).sum())

        assert grads[k][0] == pytest.approx(J(sh.velocity), rel=1e-12)
        g = grads[k][1].cpu().numpy()
        rng = np.random.default_rng(k)
        for _ in range(3):
            cell = (int(rng.integers(1, sh.shape[0] - 1)), int(rng.integers(1, sh.shape[1] - 1)))
            hs = 1e-4 * sh.velocity[cell]
            vp, vm = sh.velocity.copy(), sh.velocity.copy()
            vp[cell] += hs
            vm[cell] -= hs
            fd = (J(vp) - J(vm)) / (2 * hs)
            assert fd == pytest.approx(g[cell], rel=1e-5, abs=1e-7 * np.abs(g).max())


@pytest.mark.parametrize("order", [4, 8])
def test_plain_autograd_baseline_computes_the_same_gradient(order):
    """The plain per-shot baseline of the uneven workload (slice-form torch step + autograd) and
    the hand-written adjoint give the same gradient up to fp32 rounding (they differ in operation
    order, not in terms)."""
    from neuralmech_amd import uneven

    spec = [((60, 90), order, 70, 0, True)]
    sh = _shots(specs=spec, seed=5)[0]
    ref = _plan_reference(sh, torch.float32)
    _, g = uneven.plain_shot_gradient(sh)
    err = float((g.double() - ref["grad"]).norm() / ref["grad"].norm())
    print(f"\nplain autograd vs adjoint (order {order}): {err:.2e}")
    assert err < 1e-4
