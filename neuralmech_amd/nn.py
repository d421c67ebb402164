"""``cuwave.nn.Generator`` as ``projects/16_neuralopt/neuralfwi.py:69-71,104-124`` uses it:
``Generator(channels, sim.Nx_padded, kernel=5, output_bias=3.0)``, called with no input, returns a
``(1, 1, *Nx_padded)`` field in ``(0, 1)`` that the driver maps to the indicator
``gamma = gamma_min + (1 - gamma_min) * field`` and back-propagates the misfit gradient through.

A neural reparametrisation of the design (Hoyer, Sohl-Dickstein & Greydanus 2019, "Neural
reparameterization improves structural optimization"): a trainable latent image of
``channels[0]`` channels at ``1 / 2**(len(channels) - 2)`` of the grid, then per channel pair an
upsampling by 2 (bilinear), a ``kernel x kernel`` convolution and, except after the last one, a
group normalisation and tanh; the last convolution gives one channel whose bias starts at
``output_bias`` (``sigmoid(3) = 0.95``: the start is nearly homogeneous material), resized to the
grid exactly and passed through a sigmoid.  cuwave's network is not available; this is this
build's reading of the call site's arguments."""

from __future__ import annotations

import torch
from torch import nn
from torch.nn import functional as F


class Generator(nn.Module):
    def __init__(self, channels, shape, kernel: int = 5, output_bias: float = 0.0, seed: int | None = 0):
        super().__init__()
        channels = [int(c) for c in channels]
        if len(channels) < 2 or channels[-1] != 1 or min(channels) < 1:
            raise ValueError("Generator: at least two channel counts, the last one 1")
        if kernel < 1 or kernel % 2 == 0:
            raise ValueError("Generator: odd kernel size")
        self.shape = tuple(int(n) for n in shape)
        if len(self.shape) != 2:
            raise ValueError("Generator: 2D grids")
        stages = len(channels) - 1
        coarse = [max(1, -(-n // 2 ** (stages - 1))) for n in self.shape]
        with torch.random.fork_rng(devices=[]):  # every initial parameter from ``seed``, if given
            if seed is not None:
                torch.manual_seed(int(seed))
            self.latent = nn.Parameter(torch.randn(1, channels[0], *coarse))
            self.convs = nn.ModuleList(
                nn.Conv2d(ci, co, kernel, padding=kernel // 2) for ci, co in zip(channels[:-1], channels[1:]))
        self.norms = nn.ModuleList(nn.GroupNorm(1, co) for co in channels[1:-1])
        with torch.no_grad():
            self.convs[-1].bias.fill_(float(output_bias))

    def forward(self) -> torch.Tensor:
        x = self.latent
        for i, conv in enumerate(self.convs):
            if i > 0:
                x = F.interpolate(x, scale_factor=2, mode="bilinear", align_corners=False)
            x = conv(x)
            if i < len(self.norms):
                x = torch.tanh(self.norms[i](x))
        x = F.interpolate(x, size=self.shape, mode="bilinear", align_corners=False)
        return torch.sigmoid(x)
