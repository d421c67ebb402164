"""Drop-ins for the step kernels of ``projects/18_simple_wave_fd`` in the reference's own layout.

Each function takes torch tensors on the ROCm device holding ``(N, M_padded)`` arrays with the
boundary ring inside (node ``(i, j)`` at ``i * M_padded + j``), the same scalar arguments as the
CUDA kernel it replaces, and launches on torch's current stream:

=====================  ================================================  ============================
function               replaces                                          storage
=====================  ================================================  ============================
``padded_step``        ``padded_step`` (``step2D_wave.cu:59-81``)        fp32, ``M_padded = ceil32``
``basic_step``         ``basic_step`` / ``coalesced_step`` (``:8-55``)   fp32, ``M_padded = M``
``shared_step``        ``shared_step`` (``:187-240``)                    fp32, LDS tile
``temporal_step``      ``temporal_step`` (``:244-304``)                  fp32, 8 steps per launch
``mixed_step``         ``mixed_step`` (``:85-124``)                      fp16
``scaled_step``        ``scaled_step`` (``:128-183``)                    int16 + fp32 scale / 32
``step1d``             ``step1D`` (``step1D_wave.cu:4-17``)              fp32
=====================  ================================================  ============================

The arithmetic is the slice form of ``fd_step`` (``2D_wave_cupy.py:54-62``) in fp32 with IEEE
division, so results are bit-identical to NumPy evaluating that function on float32 arrays.
``u2`` may be ``u0`` (the in-place loop of ``2D_wave_cuda_padding.py:83-85``).
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._lib import check


def _st(t: torch.Tensor):
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _need(t: torch.Tensor, dtype, name):
    if not t.is_cuda or t.dtype != dtype or not t.is_contiguous():
        raise TypeError(f"{name} must be a contiguous {dtype} tensor on the GPU")


def _step2d(u0, u1, u2, wavespeed, dt, dx, dy, N, M, M_padded, variant):
    for t, n in ((u0, "u0"), (u1, "u1"), (u2, "u2")):
        _need(t, torch.float32, n)
        if t.numel() < N * M_padded:
            raise ValueError(f"{n} holds fewer than N * M_padded elements")
    check(_lib.load().wfd_ref_step2d(u0.data_ptr(), u1.data_ptr(), u2.data_ptr(), float(wavespeed), float(dt),
                                     float(dx), float(dy), int(N), int(M), int(M_padded), variant, _st(u0)))
    return u2


def padded_step(u0, u1, u2, wavespeed, dt, dx, dy, N, M, M_padded):
    return _step2d(u0, u1, u2, wavespeed, dt, dx, dy, N, M, M_padded, 0)


def basic_step(u0, u1, u2, wavespeed, dt, dx, dy, N, M):
    return _step2d(u0, u1, u2, wavespeed, dt, dx, dy, N, M, M, 0)


coalesced_step = basic_step


def shared_step(u0, u1, u2, wavespeed, dt, dx, dy, N, M, M_padded):
    return _step2d(u0, u1, u2, wavespeed, dt, dx, dy, N, M, M_padded, 1)


def temporal_step(u0, u1, v0, v1, wavespeed, dt, dx, dy, N, M, M_padded, steps=8):
    """``steps`` (2, 4, 8) steps from ``(u0, u1)``; the older / newer level land in ``(v0, v1)``."""
    for t, n in ((u0, "u0"), (u1, "u1"), (v0, "v0"), (v1, "v1")):
        _need(t, torch.float32, n)
        if t.numel() < N * M_padded:
            raise ValueError(f"{n} holds fewer than N * M_padded elements")
    check(_lib.load().wfd_ref_temporal(u0.data_ptr(), u1.data_ptr(), v0.data_ptr(), v1.data_ptr(), float(wavespeed),
                                       float(dt), float(dx), float(dy), int(N), int(M), int(M_padded), int(steps),
                                       _st(u0)))


def mixed_step(u0, u1, u2, wavespeed, dt, dx, dy, N, M, M_padded):
    for t, n in ((u0, "u0"), (u1, "u1"), (u2, "u2")):
        _need(t, torch.float16, n)
        if t.numel() < N * M_padded:
            raise ValueError(f"{n} holds fewer than N * M_padded elements")
    check(_lib.load().wfd_ref_mixed_step(u0.data_ptr(), u1.data_ptr(), u2.data_ptr(), float(wavespeed), float(dt),
                                         float(dx), float(dy), int(N), int(M), int(M_padded), _st(u0)))
    return u2


def scaled_step(q0, s0, q1, s1, q2, s2, wavespeed, dt, dx, dy, N, M, M_padded):
    for t, n in ((q0, "q0"), (q1, "q1"), (q2, "q2")):
        _need(t, torch.int16, n)
        if t.numel() < N * M_padded:
            raise ValueError(f"{n} holds fewer than N * M_padded elements")
    for t, n in ((s0, "s0"), (s1, "s1"), (s2, "s2")):
        _need(t, torch.float32, n)
        if t.numel() < N * (M_padded // 32):
            raise ValueError(f"{n} holds fewer than N * M_padded / 32 scales")
    check(_lib.load().wfd_ref_scaled_step(q0.data_ptr(), s0.data_ptr(), q1.data_ptr(), s1.data_ptr(), q2.data_ptr(),
                                          s2.data_ptr(), float(wavespeed), float(dt), float(dx), float(dy), int(N),
                                          int(M), int(M_padded), _st(q0)))


def step1d(u0, u1, u2, wavespeed, dt, dx, N):
    for t, n in ((u0, "u0"), (u1, "u1"), (u2, "u2")):
        _need(t, torch.float32, n)
        if t.numel() < N:
            raise ValueError(f"{n} holds fewer than N elements")
    check(_lib.load().wfd_ref_step1d(u0.data_ptr(), u1.data_ptr(), u2.data_ptr(), float(wavespeed), float(dt),
                                     float(dx), int(N), _st(u0)))
    return u2
