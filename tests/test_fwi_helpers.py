"""CPU tests of the helpers around the solver that the FWI drivers import: ``Lbfgs``
(``fwi_penalization.py:129-153``), ``TotalVariation`` and ``Projection`` (``fwi_penalization.py:158``,
``neuralfwi.py:66,104,118``), ``random_ellipses`` (``16_neuralopt/helper.py:72-79``) and the
``Generator`` ansatz (``neuralfwi.py:69``).  Each is pinned by a closed form or by an independent
formula, not by its own code."""

import math

import numpy as np
import pytest
import torch

from neuralmech_amd import geometry, wave
from neuralmech_amd.nn import Generator
from neuralmech_amd.optimization import Lbfgs
from neuralmech_amd.regularization import Projection, TotalVariation


def _quadratic(diag, b):
    A = torch.diag(torch.as_tensor(diag, dtype=torch.float64))
    b = torch.as_tensor(b, dtype=torch.float64)
    return A, b, (lambda x: float(0.5 * x @ A @ x - b @ x)), (lambda x: A @ x - b)


def test_lbfgs_two_loop_equals_explicit_bfgs_inverse_update():
    """With pairs (s_i, y_i) the two-loop direction is -H g, H built by the textbook inverse update
    H <- (I - rho s y^T) H (I - rho y s^T) + rho s s^T from H0 = (s.y / y.y) I of the newest pair
    (Nocedal & Wright eq. 6.17, 7.20)."""
    rng = np.random.default_rng(0)
    n = 6
    M = rng.standard_normal((n, n))
    A = torch.as_tensor(M @ M.T + n * np.eye(n))
    opt = Lbfgs(k=3)
    pairs = []
    for _ in range(3):
        s = torch.as_tensor(rng.standard_normal(n))
        y = A @ s
        pairs.append((s, y))
        opt.pairs.append((s, y, 1.0 / float(s @ y)))
    s, y = pairs[-1]
    H = torch.eye(n, dtype=torch.float64) * float(s @ y) / float(y @ y)
    I = torch.eye(n, dtype=torch.float64)
    for s, y in pairs:
        rho = 1.0 / float(s @ y)
        H = (I - rho * torch.outer(s, y)) @ H @ (I - rho * torch.outer(y, s)) + rho * torch.outer(s, s)
    g = torch.as_tensor(rng.standard_normal(n))
    assert torch.allclose(opt.direction(g), -H @ g, rtol=1e-12, atol=1e-12)


def test_lbfgs_first_step_and_convergence():
    A, b, f, grad = _quadratic([1.0, 10.0, 100.0, 1000.0], [1.0, 2.0, 3.0, 4.0])
    x = torch.zeros(4, dtype=torch.float64)
    opt = Lbfgs(k=4, first_step=0.05)  # too long along the stiff axis: Armijo halves it twice
    g = grad(x)
    x1, f1, alpha = opt.search(x, g, f(x), f)
    assert alpha == 0.25 and float((x1 - x).abs().max()) == pytest.approx(0.0125) and f1 < f(x)
    x = x1
    for _ in range(40):
        x, _, _ = opt.search(x, grad(x), f(x), f)
    assert torch.allclose(x, torch.linalg.solve(A, b), rtol=1e-8, atol=1e-10)


def test_lbfgs_rosenbrock():
    f = lambda x: float((1 - x[0]) ** 2 + 100 * (x[1] - x[0] ** 2) ** 2)
    grad = lambda x: torch.stack([-2 * (1 - x[0]) - 400 * x[0] * (x[1] - x[0] ** 2), 200 * (x[1] - x[0] ** 2)])
    x = torch.tensor([-1.2, 1.0], dtype=torch.float64)
    opt = Lbfgs(k=5, first_step=0.1)
    for _ in range(200):
        x, _, _ = opt.search(x, grad(x), f(x), f)
    assert torch.allclose(x, torch.ones(2, dtype=torch.float64), atol=1e-5)


def test_lbfgs_projection_and_give_up():
    A, b, f, grad = _quadratic([1.0, 1.0], [5.0, -5.0])  # minimum at (5, -5), outside the box
    clip = lambda v: v.clamp(0.0, 1.0)
    x = torch.full((2,), 0.5, dtype=torch.float64)
    opt = Lbfgs(k=2, first_step=0.2)
    for _ in range(20):
        x, _, _ = opt.search(x, grad(x), f(x), f, clip)
    assert torch.allclose(x, torch.tensor([1.0, 0.0], dtype=torch.float64))
    # an objective that rises in every direction: the search gives up and keeps x
    up = lambda v: float(((v - 0.5) ** 2).sum()) + 1.0
    x0 = torch.full((2,), 0.5, dtype=torch.float64)
    xn, fn, alpha = Lbfgs(first_step=0.1, min_alpha=1e-2).search(x0, torch.tensor([1.0, -1.0], dtype=torch.float64),
                                                                1.0, up)
    assert alpha == 0.0 and torch.equal(xn, x0) and fn == 1.0
    with pytest.raises(ValueError):
        Lbfgs(shrink=1.5)


def _tanh_disc(n, R, w):
    h = 1.0 / (n - 3)
    x = (torch.arange(n, dtype=torch.float64) - 1) * h
    X, Y = torch.meshgrid(x, x, indexing="ij")
    r = torch.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2)
    return 0.5 * (1 + torch.tanh((r - R) / w)), ((n - 1) * h) ** 2


def test_total_variation_closed_forms():
    # a disc with a narrow tanh transition: int |grad g| = 2 pi R (penalization.py:56-57)
    g, area = _tanh_disc(515, 0.2, 0.01)
    assert float(TotalVariation(1.0, eps=0.0)(g)) == pytest.approx(2 * math.pi * 0.2, rel=1e-4)
    tv = float(TotalVariation(1.0, eps=1e-2)(g))
    assert 2 * math.pi * 0.2 < tv < 2 * math.pi * 0.2 + 1e-2 * area
    # a ramp g = a x + c y: |grad g| = |(a, c)| exactly on every cell; weight scales linearly
    n = 40
    h = 1.0 / (n - 3)
    x = torch.arange(n, dtype=torch.float64) * h
    X, Y = torch.meshgrid(x, x, indexing="ij")
    ramp = 3.0 * X - 4.0 * Y
    assert float(TotalVariation(2.0, eps=0.0)(ramp)) == pytest.approx(2.0 * 5.0 * ((n - 1) * h) ** 2, rel=1e-12)
    # a constant: eps |Omega|
    assert float(TotalVariation(1.0, eps=0.3)(torch.ones(n, n))) == pytest.approx(0.3 * ((n - 1) * h) ** 2, rel=1e-6)
    # explicit spacing (0.5, 0.25): node differences 3h and -4h give the cell gradient h (6, -16)
    assert float(TotalVariation(1.0, eps=0.0, spacing=(0.5, 0.25))(ramp)) == pytest.approx(
        h * math.sqrt(36 + 256) * (n - 1) ** 2 * 0.125, rel=1e-12)


def test_total_variation_gradient_matches_central_differences():
    rng = np.random.default_rng(3)
    g = torch.as_tensor(rng.uniform(0, 1, (12, 15)))
    tv = TotalVariation(0.7, eps=0.05)
    grad = tv.grad(g)
    assert grad.shape == g.shape and grad.dtype == g.dtype
    d = 1e-6
    for i, j in [(0, 0), (5, 7), (11, 14), (3, 0), (11, 2)]:
        e = torch.zeros_like(g)
        e[i, j] = d
        fd = (float(tv(g + e)) - float(tv(g - e))) / (2 * d)
        assert float(grad[i, j]) == pytest.approx(fd, rel=1e-6, abs=1e-9)
    g32 = g.float()
    assert tv.grad(g32).dtype == torch.float32


def test_projection_closed_forms_and_chain():
    p = Projection(3.0, 0.5)
    x = torch.tensor([0.0, 0.5, 1.0], dtype=torch.float64)
    assert torch.allclose(p(x), torch.tensor([0.0, 0.5, 1.0], dtype=torch.float64), atol=1e-15)
    q = Projection(8.0, 0.3)
    a, b = math.tanh(8 * 0.3), math.tanh(8 * 0.7)
    assert float(q(torch.tensor(0.3, dtype=torch.float64))) == pytest.approx(a / (a + b), rel=1e-14)
    assert float(q(torch.tensor(0.0, dtype=torch.float64))) == pytest.approx(0.0, abs=1e-15)
    assert float(q(torch.tensor(1.0, dtype=torch.float64))) == pytest.approx(1.0, rel=1e-15)
    xs = torch.linspace(0.05, 0.95, 7, dtype=torch.float64)
    d = 1e-6
    fd = (q(xs + d) - q(xs - d)) / (2 * d)
    assert torch.allclose(q.derivative(xs), fd, rtol=1e-7)
    gout = torch.linspace(-1, 1, 7, dtype=torch.float64)
    assert torch.allclose(q.grad(xs, gout), gout * fd, rtol=1e-7, atol=1e-12)
    with pytest.raises(ValueError):
        Projection(0.0)


def test_random_ellipses():
    from scipy import ndimage

    n = 200
    coords = wave.grid_coords((n, n), (1.0 / (n - 3), 1.0 / (n - 3)))
    bounds = [(0.15, 0.85), (0.1, 0.8)]
    m = geometry.random_ellipses(coords, 4, (0.03, 0.08), bounds, overlap=False, rng=np.random.default_rng(5))
    _, count = ndimage.label(m)
    assert count == 4
    x, y = coords[0][m], coords[1][m]
    assert x.min() >= 0.15 - 0.08 and x.max() <= 0.85 + 0.08 and y.min() >= 0.1 - 0.08 and y.max() <= 0.8 + 0.08
    # equal semi-axes: a disc of area pi r^2 (node count times the cell area)
    d = geometry.random_ellipses(coords, 1, (0.2, 0.2), [(0.5, 0.5), (0.5, 0.5)], rng=np.random.default_rng(0))
    h = 1.0 / (n - 3)
    assert d.sum() * h * h == pytest.approx(math.pi * 0.04, rel=0.01)
    same = geometry.random_ellipses(coords, 4, (0.03, 0.08), bounds, overlap=False, rng=np.random.default_rng(5))
    assert np.array_equal(m, same)
    with pytest.raises(RuntimeError):
        geometry.random_ellipses(coords, 50, (0.2, 0.25), bounds, overlap=False, rng=np.random.default_rng(0),
                                 max_tries=200)


def test_generator_ansatz():
    g = Generator([16, 16, 16, 16, 8, 1], (67, 70), kernel=5, output_bias=3.0)
    y = g()
    assert y.shape == (1, 1, 67, 70) and float(y.detach().min()) > 0 and float(y.detach().max()) < 1
    assert abs(float(y.detach().mean()) - 1 / (1 + math.exp(-3.0))) < 0.05  # nearly homogeneous start
    y.backward(torch.ones_like(y))
    assert all(p.grad is not None and float(p.grad.abs().sum()) > 0 for p in g.parameters())
    assert torch.equal(Generator([4, 1], (9, 9), seed=1)(), Generator([4, 1], (9, 9), seed=1)())
    with pytest.raises(ValueError):
        Generator([4, 2], (9, 9))
