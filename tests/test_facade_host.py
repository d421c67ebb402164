"""CPU tests of the cuwave-style facade: descriptor validation (CFL guard), coordinates, source
snapping and scaling, sponge padding, resampling and wavelets.  No kernel is launched."""

import math
from dataclasses import replace

import numpy as np
import pytest
import torch

from neuralmech_amd import boundary, media, signals, utils, wave
from neuralmech_amd.scalar import ScalarWave
from oracle import slice_ref


def _sim(**kw):
    base = dict(Nx=(64, 80), dx=(0.01, 0.01), N=100, dt=0.004, precision="float32", space_order=4, wavespeed=1.0,
                device="cpu")
    base.update(kw)
    return ScalarWave(**base)


def test_cfl_guard_rejects_unstable_dt():
    dt_max = wave.stable_dt((0.01, 0.01), 1.0, 4)
    assert dt_max == pytest.approx(slice_ref.stable_dt((0.01, 0.01), 1.0, 4))
    _sim(dt=0.999 * dt_max)
    with pytest.raises(ValueError):
        _sim(dt=1.01 * dt_max)
    v = np.ones((64, 80))
    v[10, 10] = 3.0  # the fastest node sets the limit
    with pytest.raises(ValueError):
        _sim(dt=0.9 * dt_max, wavespeed=v)


def test_descriptor_basics_and_replace():
    s = _sim()
    assert s.Nx_padded == (64, 80) and s.ndim == 2 and s.cell_volume == pytest.approx(1e-4)
    prof = boundary.sponge(s, None, 10, 0.2)
    s2 = replace(s, damping=prof)
    assert s2.damping is prof and s2.Nx == s.Nx
    with pytest.raises(ValueError):
        _sim(space_order=5)


def test_grid_coords_convention():
    x, y = wave.grid_coords((6, 7), (0.5, 0.25))
    assert x[1, 0] == 0.0 and y[0, 1] == 0.0 and x[0, 0] == -0.5 and y[0, 6] == 1.25
    assert x.shape == (6, 7)


def test_point_source_snaps_and_scales():
    s = _sim()
    src = utils.point_source(s, (0.203, 0.5), np.ones(100))
    assert src.idx.shape == (2, 1) and tuple(src.idx[:, 0]) == (21, 51)
    assert np.allclose(src.signal, 1.0 / 1e-4)
    far = utils.point_source(s, (-5.0, 99.0), np.ones(100))
    assert tuple(far.idx[:, 0]) == (1, 78)  # clipped to the interior
    many = utils.shots(s, utils.line((0.1, 0.0), (0.3, 0.0), 5), np.ones(100))
    assert len(many) == 5 and [int(m.idx[0, 0]) for m in many] == [11, 16, 21, 26, 31]
    sens = utils.Sensors(s, utils.line((0.0, 0.1), (0.0, 0.5), 3))
    assert sens.nodes.shape == (2, 3) and sens.count == 3


def test_sponge_profiles_and_padding():
    s = _sim()
    prof = boundary.sponge(s, None, 12, 0.3, faces=[(0, "lo"), (1, "hi")])
    assert np.array_equal(prof[0], slice_ref.sponge_profile(64, 12, 0.3, lo=True, hi=False))
    assert np.array_equal(prof[1], slice_ref.sponge_profile(80, 12, 0.3, lo=False, hi=True))
    assert np.array_equal(media.sponge_profile(50, 7, 0.1), slice_ref.sponge_profile(50, 7, 0.1))
    Nx, width, origin, region = boundary.pad_for_sponge((100, 60), (0.01, 0.01), 0.1, faces=[(0, 0), (0, 1), (1, 1)])
    assert width == 10 and Nx == (120, 70) and origin == (-0.1, 0.0)
    assert region == (slice(10, 110), slice(0, 60))


def test_resample_is_linear_interpolation():
    t = np.arange(50) * 0.1
    rec = np.stack([2 * t + 1, -t], axis=1)
    out = utils.resample(rec, 0.1, 0.025, 150)
    tt = np.arange(150) * 0.025
    inside = tt <= t[-1]
    assert np.allclose(out[inside, 0], 2 * tt[inside] + 1) and np.allclose(out[inside, 1], -tt[inside])
    assert np.array_equal(utils.resample(rec, 0.1, 0.1, 50), rec)


def test_sineburst_window():
    t = np.linspace(0, 1, 1001)
    s = signals.sineburst(t, 2.0, 10.0, 3)
    assert np.all(s[t > 0.3 + 1e-9] == 0) and s[0] == 0
    assert np.max(np.abs(s)) <= 2.0 and np.max(np.abs(s)) > 1.5
    r = signals.ricker(np.array([0.1]), 10.0)
    assert r[0] == 1.0


def test_density_indicator_parametrisation():
    s = _sim()
    assert wave._unit_indicator(None)
    assert not wave._unit_indicator(np.ones(s.Nx)) and not wave._unit_indicator(np.full(s.Nx, 0.5))
    g = torch.full(s.Nx, 0.25, dtype=torch.float64)
    m, k = s.material(g)
    v = s.velocity()
    rho = float(torch.as_tensor(s.density, dtype=torch.float64))
    assert torch.allclose(m, 0.25 * rho * torch.ones_like(m))
    assert torch.allclose(k, 0.25 * rho * v ** 2)
    dm, dk = s.parametrization_jacobian(g)
    assert torch.allclose(dm, torch.full_like(g, rho)) and torch.allclose(dk, rho * v ** 2)
    with pytest.raises(ValueError):
        s.material(np.ones((3, 3)))


def test_save_csv_writes_the_reference_columns(tmp_path):
    from neuralmech_amd import harness

    harness.save_csv(tmp_path / "wave_scaling_0.csv", x=[1, 4, 9], y=[1e-6, 2e-6, 3e-6])
    lines = (tmp_path / "wave_scaling_0.csv").read_text().splitlines()
    assert lines[0] == "x y" and lines[1].split() == ["1", "1e-06"] and len(lines) == 4
    import pytest

    with pytest.raises(ValueError):
        harness.save_csv(tmp_path / "bad.csv", x=[1, 2], y=[1])


def test_fwi_driver_helpers():
    from neuralmech_amd import evals, geometry, optimization

    truth = np.ones((20, 20))
    truth[5:9, 5:9] = 1e-4
    g = truth.copy()
    assert evals.l2_error(g, truth) == 0.0 and evals.f1_score(g, truth) == 1.0
    g2 = np.ones((20, 20))
    assert evals.f1_score(g2, truth) == 0.0 and evals.l2_error(g2, truth) > 0
    # Adam on f(x) = |x - 3|^2 / 2 converges; the first step has size lr in every component
    opt = optimization.Adam(lr=0.1)
    x = torch.zeros(5, dtype=torch.float64)
    x1 = opt.step(x, x - 3.0)
    assert torch.allclose(x1, torch.full_like(x, 0.1))
    for _ in range(400):
        x1 = opt.step(x1, x1 - 3.0)
    assert float((x1 - 3.0).abs().max()) < 0.05
    coords = wave.grid_coords((64, 64), (1 / 61, 1 / 61))
    m = geometry.stacked_circles(coords, 4, 0.05, (0.1, 0.5), 0.5, axis=1, ratio=0.5, order="ascending")
    assert m.shape == (64, 64) and m.sum() > 0
    y = coords[1][m]
    assert y.min() >= 0.1 - 0.05 and y.max() <= 0.5 + 0.05
    assert torch.equal(utils.threshold(torch.tensor([0.2, 0.5, 0.9])), torch.tensor([0.0, 1.0, 1.0]))


def test_gradient_plan_cache_leases(monkeypatch):
    """The kept gradient plan is reused only when no forward holds it; a second outstanding
    forward gets a plan of its own; releasing (or dropping) a lease frees the kept plan."""
    from neuralmech_amd import autograd, solver

    made = []

    class FakePlan:
        def __init__(self, grid, nsteps, seg=0, max_bytes=0):
            made.append(self)

    monkeypatch.setattr(autograd, "GradientPlan", FakePlan)
    autograd.clear_plans()
    g = solver.WaveGrid((20, 30), (1.0, 1.0), 0.1, 4, torch.float32, device="cpu")
    p1 = autograd._plan_for(g, 50, 0, 0)
    lease1 = autograd._Lease(p1)
    p2 = autograd._plan_for(g, 50, 0, 0)
    assert p2 is not p1 and len(made) == 2  # kept plan leased: a fresh one
    lease1.release()
    assert autograd._plan_for(g, 50, 0, 0) is p1 and len(made) == 2
    lease = autograd._Lease(p1)
    del lease  # a dropped graph releases its lease
    assert autograd._plan_for(g, 50, 0, 0) is p1
    g2 = solver.WaveGrid((20, 31), (1.0, 1.0), 0.1, 4, torch.float32, device="cpu")
    p3 = autograd._plan_for(g2, 50, 0, 0)
    assert p3 is not p1 and autograd._plan_for(g, 50, 0, 0) is not p1  # one plan kept at a time
    autograd.clear_plans()
