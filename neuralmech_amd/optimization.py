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
