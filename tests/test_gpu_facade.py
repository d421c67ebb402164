"""GPU tests of the cuwave-style facade against the C oracle and the solver objects."""

from dataclasses import replace

import numpy as np
import pytest

from oracle import c_oracle, slice_ref
from tests.wave_cases import rel_l2

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _setup(order=8, precision="float32", n=(96, 128), N=120, seed=0):
    from neuralmech_amd import boundary, utils, wave
    from neuralmech_amd.scalar import ScalarWave

    rng = np.random.default_rng(seed)
    h = (10.0, 10.0)
    v = rng.uniform(1500.0, 3000.0, n)
    dt = 0.8 * wave.stable_dt(h, v, order)
    sim = ScalarWave(n, h, N, dt, precision=precision, space_order=order, wavespeed=v)
    sim = replace(sim, damping=boundary.sponge(sim, None, 12, 0.08))
    t = np.arange(N) * dt
    sources = utils.shots(sim, utils.line((100.0, 60.0), (1100.0, 60.0), 3), slice_ref.ricker(t, 15.0))
    sensors = utils.Sensors(sim, utils.line((0.0, 50.0), (1200.0, 50.0), 20))
    return sim, sources, sensors, v


def test_simulate_matches_oracle_bitexact():
    from neuralmech_amd import wave

    sim, sources, sensors, v = _setup()
    u, traces = wave.simulate(sim, sources[1], None, sensors)
    torch.cuda.synchronize()
    kc = ((torch.as_tensor(v) * sim.dt) ** 2).float().numpy()
    src = [tuple(sources[1].idx[:, 0])]
    rec = [tuple(p) for p in sensors.nodes.T]
    tr_o, _, uc_o = c_oracle.forward(kc, sim.space_order, sim.dx, sim.N, beta_profiles=sim.damping, src_nodes=src,
                                     src_sig=sources[1].signal.astype(np.float32), rec_nodes=rec)
    assert np.array_equal(u.cpu().numpy(), uc_o)
    assert np.array_equal(traces.cpu().numpy(), tr_o)


def test_measure_batches_shots_identically():
    from neuralmech_amd import utils, wave

    sim, sources, sensors, _ = _setup()
    recs = utils.measure(sim, sources, None, sensors)
    for s, r in zip(sources, recs):
        _, tr = wave.simulate(sim, s, None, sensors)
        assert torch.equal(r, tr)


def test_simulate_snapshots():
    from neuralmech_amd import wave

    sim, sources, sensors, _ = _setup(N=60)
    u, tr, frames = wave.simulate(sim, sources[0], None, sensors, record_every=20)
    assert len(frames) == 3 and np.array_equal(frames[-1], u.cpu().numpy())
    u2 = wave.simulate(replace(sim, N=20), sources[0])
    assert np.array_equal(frames[0], u2.cpu().numpy())


def test_wavespeed_misfit_gradient_fp64_matches_central_differences():
    from neuralmech_amd import utils

    sim, sources, sensors, v = _setup(order=4, precision="float64", n=(48, 56), N=80)
    rng = np.random.default_rng(1)
    obs = [rng.standard_normal((80, sensors.count)) * 1e-3 for _ in sources]
    J, grad = utils.wavespeed_misfit_gradient(sim, sources, v, sensors, obs)
    assert J == pytest.approx(utils.wavespeed_misfit(sim, sources, v, sensors, obs), rel=1e-12)
    g = grad.cpu().numpy()
    for cell in [(10, 20), (25, 30), (40, 12)]:
        hstep = 1e-4 * v[cell]
        vp, vm = v.copy(), v.copy()
        vp[cell] += hstep
        vm[cell] -= hstep
        fd = (utils.wavespeed_misfit(sim, sources, vp, sensors, obs)
              - utils.wavespeed_misfit(sim, sources, vm, sensors, obs)) / (2 * hstep)
        assert fd == pytest.approx(g[cell], rel=1e-5, abs=1e-9 * np.abs(g).max())


@pytest.mark.parametrize("order", [2, 4, 8])
def test_density_indicator_misfit_gradient_fp64(order):
    """cuwave's misfit_gradient(sim, sources, gamma, sensors, observed) (2D_fwi.py:40, orders 4 and
    8 at its call sites): gradient in the density indicator by the reconstruction adjoint, against
    central differences; a unit indicator reproduces the constant-density run."""
    from neuralmech_amd import acoustic, boundary, utils, wave
    from neuralmech_amd.scalar import ScalarWave

    n, h, N = (44, 52), (1e-3, 1e-3), 150
    dt = 0.7 * wave.stable_dt(h, 6000.0, order)
    sim = ScalarWave(n, h, N, dt, precision="float64", space_order=order, wavespeed=6000.0, density=2700.0)
    sim = replace(sim, damping=boundary.sponge(sim, None, 6, 0.1))
    t = np.arange(N) * dt
    from neuralmech_amd import signals

    sources = utils.shots(sim, utils.line((8e-3, 8e-3), (30e-3, 8e-3), 3), signals.sineburst(t, 1e8, 6e5, 2))
    sensors = utils.Sensors(sim, utils.line((6e-3, 35e-3), (36e-3, 35e-3), 12))
    plain = utils.measure(sim, sources, None, sensors)
    ind = acoustic.measure_indicator(sim, sources, torch.ones(n, dtype=torch.float64), sensors)
    for a, b in zip(plain, ind):
        assert rel_l2(b.cpu().numpy(), a.cpu().numpy()) < 1e-11
    rng = np.random.default_rng(4)
    gamma = torch.as_tensor(rng.uniform(0.5, 1.0, n))
    truth = torch.as_tensor(rng.uniform(0.5, 1.0, n))
    obs = [r.cpu().numpy() for r in utils.measure(sim, sources, truth, sensors)]
    J, grad = utils.misfit_gradient(sim, sources, gamma, sensors, obs)
    assert J == pytest.approx(utils.misfit(sim, sources, gamma, sensors, obs), rel=1e-12)
    assert J > 0 and torch.isfinite(grad).all()
    for cell in [(20, 26), (15, 30), (28, 18)]:
        eps = 1e-5
        gp, gm = gamma.clone(), gamma.clone()
        gp[cell] += eps
        gm[cell] -= eps
        fd = (utils.misfit(sim, sources, gp, sensors, obs) - utils.misfit(sim, sources, gm, sensors, obs)) / (2 * eps)
        assert fd == pytest.approx(float(grad[cell]), rel=1e-5, abs=1e-6 * float(grad.abs().max()))


def _plate_problem(Nx=(96, 96), n_elements=8):
    """projects/18_fwi/plate.py at a reduced size: a void measured at order 8 with a density
    indicator, resampled onto the order-4 inversion grid (every other element fires)."""
    import math

    from neuralmech_amd import geometry, signals, utils, wave
    from neuralmech_amd.scalar import ScalarWave

    L, T = (1.0, 1.0), 1.6
    dx = tuple(l / (n - 3) for l, n in zip(L, Nx))
    elements = utils.line((0.1, L[1]), (0.9, L[1]), n_elements)

    def simulation(order):
        N = math.ceil(T / (0.9 * wave.stable_dt(dx, 1.0, order))) + 1
        sim = ScalarWave(Nx, dx, N, T / (N - 1), (4, 64), precision="float32", space_order=order, wavespeed=1.0,
                         density=1.0)
        t = np.linspace(0, T, N)
        return sim, t, utils.shots(sim, elements, signals.sineburst(t, 1.0, 8.0, 2)), utils.Sensors(sim, elements)

    sim_obs, t_obs, src_obs, sens_obs = simulation(8)
    sim, t, src, sens = simulation(4)
    void = geometry.circle(wave.grid_coords(Nx, dx), (0.5, 0.45), 0.12)
    truth = np.where(void, 1e-4, 1.0)
    observed = utils.measure(sim_obs, src_obs[::2], truth, sens_obs)
    observed = [utils.resample(r, t_obs[1], t[1], sim.N) for r in observed]
    return sim, src[::2], sens, observed


def test_plate_fwi_flow_with_density_indicator():
    """projects/18_fwi/2D_fwi.py at a reduced size: one indicator gradient of all shots; the
    steepest-descent step lowers the misfit."""
    from neuralmech_amd import utils

    sim, src, sens, observed = _plate_problem()
    gamma = torch.ones(sim.Nx, dtype=torch.float64, device="cuda")
    J0, g = utils.misfit_gradient(sim, src, gamma, sens, observed)
    assert J0 > 0 and torch.isfinite(g).all() and float(g.abs().max()) > 0
    step = 0.05 / float(g.abs().max())
    J1 = utils.misfit(sim, src, (gamma - step * g).clamp(1e-3, 1.0), sens, observed)
    assert J1 < J0


def test_penalized_lbfgs_flow():
    """projects/9_structural_optimization/fwi_penalization.py:125-153 at a reduced size: L-BFGS on
    misfit + total variation with the box projection; every accepted step lowers the penalised
    objective (Armijo) and the iterate stays in [gamma_min, 1]."""
    from neuralmech_amd import utils
    from neuralmech_amd.optimization import Lbfgs
    from neuralmech_amd.regularization import TotalVariation

    sim, src, sens, observed = _plate_problem()
    penalty = TotalVariation(3e-3, eps=1e-2)
    gamma = torch.ones(sim.Nx_padded, dtype=torch.float32, device="cuda")
    opt = Lbfgs(k=4, first_step=0.05, armijo=1e-4, shrink=0.5, min_alpha=1e-3)
    objective = lambda g: float(utils.misfit(sim, src, g, sens, observed)) + float(penalty(g))
    history = []
    for _ in range(4):
        cost, gradient = utils.misfit_gradient(sim, src, gamma, sens, observed)
        gradient = gradient + penalty.grad(gamma)
        total = float(cost) + float(penalty(gamma))
        history.append(total)
        gamma, f_new, alpha = opt.search(gamma, gradient, total, objective, lambda g: g.clamp(1e-3, 1.0))
        assert alpha > 0 and f_new < total
        assert float(gamma.min()) >= 1e-3 and float(gamma.max()) <= 1.0 and gamma.dtype == torch.float32
    assert history[-1] < history[0]


def test_neural_and_projection_ansatz_descend():
    """projects/16_neuralopt/neuralfwi.py:99-124 at a reduced size: the misfit gradient chained
    through the Generator (field.backward) and through the Heaviside projection gives a descent
    direction for the network weights and for the projected design variables."""
    from neuralmech_amd import utils
    from neuralmech_amd.nn import Generator
    from neuralmech_amd.regularization import Projection

    gmin = 1e-3
    sim, src, sens, observed = _plate_problem()
    model = Generator([8, 8, 8, 4, 1], sim.Nx_padded, kernel=5, output_bias=3.0).to("cuda")
    field = model()
    gamma = gmin + (1.0 - gmin) * field.detach()[0, 0]
    J0, gradient = utils.misfit_gradient(sim, src, gamma, sens, observed)
    field.backward(torch.as_tensor((1.0 - gmin) * gradient).to(field.dtype)[None, None])
    grads = [p.grad.detach().clone() for p in model.parameters()]
    assert all(torch.isfinite(g).all() for g in grads) and sum(float(g.abs().sum()) for g in grads) > 0
    norm = max(float(g.abs().max()) for g in grads)
    with torch.no_grad():
        for p, g in zip(model.parameters(), grads):
            p -= (0.02 / norm) * g
        gamma1 = gmin + (1.0 - gmin) * model()[0, 0]
    assert utils.misfit(sim, src, gamma1, sens, observed) < J0

    proj = Projection(3.0, 0.5)
    x = torch.full(sim.Nx_padded, 0.9, dtype=torch.float32, device="cuda")
    J0, gradient = utils.misfit_gradient(sim, src, gmin + (1.0 - gmin) * proj(x), sens, observed)
    gx = proj.grad(x, (1.0 - gmin) * gradient)
    x1 = (x - (0.05 / float(gx.abs().max())) * gx).clamp(0.0, 1.0)
    assert utils.misfit(sim, src, gmin + (1.0 - gmin) * proj(x1), sens, observed) < J0
