"""``cuwave.optimization.Adam`` as the FWI drivers use it: ``optimizer = Adam(lr=LR)``, then
``x = optimizer.step(x, gradient)`` per iteration (``projects/18_fwi/2D_fwi.py:35,41``,
``16_neuralopt/neuralfwi.py:82,116``).  The standard Adam update (Kingma & Ba, 2015) on a plain
array, state kept by the optimizer; the defaults are the paper's."""

from __future__ import annotations

import torch


class Adam:
    def __init__(self, lr: float = 1e-3, betas=(0.9, 0.999), eps: float = 1e-8):
        self.lr, self.b1, self.b2, self.eps = float(lr), float(betas[0]), float(betas[1]), float(eps)
        self.m = self.v = None
        self.t = 0

    def step(self, x, gradient):
        """One step on ``x`` for ``gradient``; returns the new array (the input is not modified)."""
        x = torch.as_tensor(x)
        g = torch.as_tensor(gradient).to(device=x.device, dtype=x.dtype)
        if self.m is None:
            self.m = torch.zeros_like(x)
            self.v = torch.zeros_like(x)
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        return x - self.lr * mh / (torch.sqrt(vh) + self.eps)


class Lbfgs:
    """``cuwave.optimization.Lbfgs`` as ``projects/9_structural_optimization/fwi_penalization.py:129-153``
    calls it: ``Lbfgs(k, first_step, armijo, shrink, min_alpha)``, then per iteration
    ``x = optimizer.search(x, gradient, cost, objective, project)[0]``.

    Limited-memory BFGS (Nocedal & Wright, Numerical Optimization, 2nd ed., Alg. 7.4 two-loop
    recursion, initial matrix ``(s.y / y.y) I`` from the newest pair, eq. 7.20) with a backtracking
    Armijo line search (Alg. 3.1) on the projected path.  The curvature pair of a call is made from
    the previous call's point and gradient (``s = x - x_prev``, ``y = g - g_prev``), so the driver
    evaluates each gradient once; a pair with ``s.y`` not positive is dropped.  Without a pair the
    direction is steepest descent scaled so that its largest entry moves by ``first_step``.  The
    trial ``project(x + alpha d)`` is accepted once ``objective(trial) <= cost + armijo * g.(trial - x)``;
    ``alpha`` shrinks by ``shrink`` per rejection and the search gives up (returns ``x`` unchanged and
    forgets the pairs) below ``min_alpha``.  cuwave's own implementation is not available; this is
    the published algorithm with the call site's argument meanings."""

    def __init__(self, k: int = 4, first_step: float = 1.0, armijo: float = 1e-4, shrink: float = 0.5,
                 min_alpha: float = 1e-3):
        if int(k) < 1 or not (0 < shrink < 1) or not (0 < min_alpha <= 1) or first_step <= 0:
            raise ValueError("Lbfgs: k >= 1, 0 < shrink < 1, 0 < min_alpha <= 1, first_step > 0")
        self.k, self.first_step, self.armijo = int(k), float(first_step), float(armijo)
        self.shrink, self.min_alpha = float(shrink), float(min_alpha)
        self.pairs: list[tuple[torch.Tensor, torch.Tensor, float]] = []  # (s, y, 1 / y.s), oldest first
        self.x_prev = self.g_prev = None

    @staticmethod
    def _dot(a, b) -> float:
        return float(torch.sum(a * b))

    def direction(self, g: torch.Tensor) -> torch.Tensor:
        """``-H g`` by the two-loop recursion over the stored pairs."""
        if not self.pairs:
            gmax = float(g.abs().max())
            return -self.first_step * g / gmax if gmax > 0 else torch.zeros_like(g)
        q = g.clone()
        alphas = []
        for s, y, rho in reversed(self.pairs):
            a = rho * self._dot(s, q)
            q -= a * y
            alphas.append(a)
        s, y, _ = self.pairs[-1]
        r = q * (self._dot(s, y) / self._dot(y, y))
        for (s, y, rho), a in zip(self.pairs, reversed(alphas)):
            b = rho * self._dot(y, r)
            r += (a - b) * s
        return -r

    def search(self, x, gradient, cost, objective, project=None):
        """One L-BFGS step from ``x`` (gradient ``gradient``, objective value ``cost``).  Returns
        ``(x_new, cost_new, alpha)``; ``alpha = 0`` when the line search gave up."""
        x = torch.as_tensor(x)
        g = torch.as_tensor(gradient).to(device=x.device, dtype=x.dtype)
        cost = float(cost)
        project = project if project is not None else (lambda v: v)
        if self.x_prev is not None:
            s, y = x - self.x_prev, g - self.g_prev
            sy = self._dot(s, y)
            if sy > 1e-12 * max(self._dot(s, s), 1e-300) ** 0.5 * max(self._dot(y, y), 1e-300) ** 0.5:
                self.pairs.append((s, y, 1.0 / sy))
                if len(self.pairs) > self.k:
                    self.pairs.pop(0)
        d = self.direction(g)
        if self._dot(g, d) >= 0:  # not a descent direction: restart from steepest descent
            self.pairs.clear()
            d = self.direction(g)
        self.x_prev, self.g_prev = x, g
        alpha = 1.0
        while alpha >= self.min_alpha:
            trial = torch.as_tensor(project(x + alpha * d)).to(device=x.device, dtype=x.dtype)
            f = float(objective(trial))
            if f <= cost + self.armijo * self._dot(g, trial - x):
                return trial, f, alpha
            alpha *= self.shrink
        self.pairs.clear()
        return x, cost, 0.0
