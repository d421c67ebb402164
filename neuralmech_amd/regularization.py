"""``cuwave.regularization`` names of the FWI drivers: ``TotalVariation(weight, eps)``
(``projects/9_structural_optimization/fwi_penalization.py:12,158``, used as ``float(penalty(g))``
and ``gradient += penalty.grad(g)`` at ``:135,146``) and ``Projection(beta, eta)``
(``projects/16_neuralopt/neuralfwi.py:13,66``, used as ``projection(x)`` and
``projection.grad(x, dJ/dprojected)`` at ``:104,118``).  cuwave's own implementations are not
available; the definitions below are stated here and pinned by closed forms in the tests."""

from __future__ import annotations

import math

import torch


def _t(a):
    return torch.as_tensor(a.get() if hasattr(a, "get") else a)


class TotalVariation:
    """Smoothed total variation ``weight * int sqrt(|grad g|^2 + eps^2) dOmega``, the form of
    ``projects/9_structural_optimization/penalization.py:47-48`` (below ``eps`` it acts like a
    Tikhonov term).  Discretisation: the gradient of each grid cell is that of the bilinear
    (trilinear) interpolant at the cell centre, i.e. the mean of the cell's edge differences along
    each axis; the integral sums the cells times the cell volume.  ``spacing`` defaults to the call
    sites' unit-plate convention ``1 / (n_d - 3)`` per axis (``fwi_penalization.py:76``).  For a
    disc of radius R with a narrow transition the value tends to ``weight * (2 pi R + eps |Omega|)``
    from below as the transition sharpens (``penalization.py:56-57``)."""

    def __init__(self, weight: float, eps: float = 1e-2, spacing=None):
        if not weight >= 0 or not eps >= 0:
            raise ValueError("TotalVariation: weight >= 0 and eps >= 0")
        self.weight, self.eps, self.spacing = float(weight), float(eps), spacing

    def _h(self, shape):
        if self.spacing is not None:
            h = [float(v) for v in (self.spacing if hasattr(self.spacing, "__len__") else [self.spacing] * len(shape))]
        else:
            h = [1.0 / (n - 3) for n in shape]
        if len(h) != len(shape) or any(not v > 0 for v in h) or any(n < 2 for n in shape):
            raise ValueError("TotalVariation: one positive spacing per axis and >= 2 nodes per axis")
        return h

    def _value(self, g: torch.Tensor) -> torch.Tensor:
        h = self._h(g.shape)
        D = g.ndim
        sq = 0.0
        for d in range(D):
            diff = torch.diff(g, dim=d) / h[d]
            for e in range(D):  # average the edge differences over the cell's other axes
                if e != d:
                    diff = 0.5 * (diff.narrow(e, 0, diff.shape[e] - 1) + diff.narrow(e, 1, diff.shape[e] - 1))
            sq = sq + diff * diff
        return self.weight * torch.sqrt(sq + self.eps ** 2).sum() * math.prod(h)

    def __call__(self, g) -> torch.Tensor:
        g = _t(g)
        return self._value(g.double()).to(g.dtype)

    def grad(self, g) -> torch.Tensor:
        """The gradient of the penalty w.r.t. every node of ``g`` (same shape, device, dtype)."""
        g = _t(g)
        x = g.detach().double().requires_grad_(True)
        with torch.enable_grad():
            (gx,) = torch.autograd.grad(self._value(x), x)
        return gx.to(g.dtype)


class Projection:
    """Smoothed Heaviside projection ``H(x) = (tanh(beta eta) + tanh(beta (x - eta))) /
    (tanh(beta eta) + tanh(beta (1 - eta)))`` (Wang, Lazarov & Sigmund 2011; the form of
    ``solvers/optimization.py:433-445``): ``H(0) = 0``, ``H(1) = 1``, sharper as ``beta`` grows.
    ``grad(x, g)`` chains a gradient ``g`` w.r.t. ``H(x)`` back to ``x``."""

    def __init__(self, beta: float, eta: float = 0.5):
        if not beta > 0 or not 0 <= eta <= 1:
            raise ValueError("Projection: beta > 0 and 0 <= eta <= 1")
        self.beta, self.eta = float(beta), float(eta)
        self._a, self._b = math.tanh(self.beta * self.eta), math.tanh(self.beta * (1.0 - self.eta))

    def __call__(self, x) -> torch.Tensor:
        x = _t(x)
        return (self._a + torch.tanh(self.beta * (x - self.eta))) / (self._a + self._b)

    def derivative(self, x) -> torch.Tensor:
        x = _t(x)
        t = torch.tanh(self.beta * (x - self.eta))
        return self.beta * (1.0 - t * t) / (self._a + self._b)

    def grad(self, x, g) -> torch.Tensor:
        x = _t(x)
        return _t(g).to(device=x.device, dtype=x.dtype) * self.derivative(x)
