"""Total focusing method over ``wfd_tfm``: the imager of ``projects/18_fwi/tfm.py`` and the analytic
signal it images (``cupyx.scipy.signal.hilbert`` at ``projects/18_fwi/2D_tfm.py:35``).

* :func:`analytic_signal` -- ``x + i H[x]`` along an axis, by the FFT (the one-sided spectrum
  doubled), as ``scipy.signal.hilbert`` computes it; torch on the device.
* :func:`tfm_kernel` -- ``tfm_kernel(h, elements, t0, dt, c, x, y)`` (``tfm.py:56-67``): element
  pairs r >= e with the pair r > e weighted 2 (reciprocity), fp32; one lane per pixel for images
  of 2^20 pixels and more, a group of up to 16 lanes per pixel below that.
* :func:`tfm` -- ``tfm(h, elements, t0, dt, c, x, y)`` (``tfm.py:39-53``): every pair, the samples
  past the last interval dropped; runs the same kernel in its all-pairs form.

``h`` is ``(E, E, Nt)`` complex (emitter, receiver, sample); ``elements`` ``(E, 2)``; ``x``, ``y``
pixel coordinates of any (equal) shape.  Images come back as fp32 tensors of that shape.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check


def analytic_signal(x, axis: int = -1) -> torch.Tensor:
    """``x + i H[x]`` along ``axis`` (``scipy.signal.hilbert``): FFT, keep DC (and Nyquist for even
    length), double the positive frequencies, drop the negative ones, inverse FFT."""
    t = torch.as_tensor(x)
    if t.is_complex():
        raise ValueError("the analytic signal of a real signal")
    n = t.shape[axis]
    spec = torch.fft.fft(t, dim=axis)
    hmul = torch.zeros(n, dtype=t.dtype, device=t.device)
    if n % 2 == 0:
        hmul[0] = hmul[n // 2] = 1.0
        hmul[1:n // 2] = 2.0
    else:
        hmul[0] = 1.0
        hmul[1:(n + 1) // 2] = 2.0
    shape = [1] * t.dim()
    shape[axis] = n
    return torch.fft.ifft(spec * hmul.view(shape), dim=axis)


def _image(h, elements, t0, dt, c, x, y, reciprocal: int):
    h = torch.as_tensor(h)
    dev = h.device if h.is_cuda else torch.device("cuda")
    h = h.to(device=dev, dtype=torch.complex64).contiguous()
    if h.dim() != 3 or h.shape[0] != h.shape[1]:
        raise ValueError("h must be (E, E, Nt)")
    E, _, Nt = h.shape
    el = torch.as_tensor(np.asarray(elements, np.float64), dtype=torch.float32, device=dev).reshape(E, 2)
    xe, ye = el[:, 0].contiguous(), el[:, 1].contiguous()
    xt = torch.as_tensor(x).to(device=dev, dtype=torch.float32)
    yt = torch.as_tensor(y).to(device=dev, dtype=torch.float32)
    if xt.shape != yt.shape:
        raise ValueError("x and y must have one shape")
    xf, yf = xt.reshape(-1).contiguous(), yt.reshape(-1).contiguous()
    image = torch.empty(xf.numel(), dtype=torch.float32, device=dev)
    st = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    check(_lib.load().wfd_tfm(h.data_ptr(), xe.data_ptr(), ye.data_ptr(), int(E), int(Nt), float(dt), float(t0),
                              float(c), xf.data_ptr(), yf.data_ptr(), int(xf.numel()), image.data_ptr(),
                              int(reciprocal), st))
    return image.reshape(xt.shape)


def tfm_kernel(h, elements, t0, dt, c, x, y) -> torch.Tensor:
    """The reference's delay-and-sum kernel form (reciprocal pairs), fp32, on the device."""
    return _image(h, elements, t0, dt, c, x, y, 1)


def tfm(h, elements, t0, dt, c, x, y) -> torch.Tensor:
    """The reference's array form (every pair, samples past ``Nt - 2`` dropped), on the device."""
    return _image(h, elements, t0, dt, c, x, y, 0)
