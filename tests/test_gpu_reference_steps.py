"""Reference-layout drop-ins against the slice-form restatement of the reference step:
bit-identical fp32 (padded, unpadded, LDS, 1D), the temporal-blocked launch bit-identical to
single steps (the reference's own invariant, README.md:39-40), fp16 and int16 block-scaled
storage bit-identical to their NumPy restatements."""

import math

import numpy as np
import pytest

from oracle import slice_ref

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _setup(NX=160, NY=320, pad=32):
    dx, dy = 1.0 / (NX - 1), 2.0 / (NY - 1)
    dt = slice_ref.reference_dt(dx, dy, 1.0, 0.95)
    Mp = -(-NY // pad) * pad
    u = np.zeros((NX, Mp), np.float32)
    u[:, :NY] = slice_ref.gaussian_ic(NX, NY, center=(0.5, 0.5), sigma=0.05)
    return dx, dy, dt, Mp, u


@pytest.mark.parametrize("variant,pad,NX,NY", [
    ("padded", 32, 160, 320),   # 4-node chunks; the last chunk holds the boundary column
    ("padded", 32, 133, 301),   # chunks straddle the boundary and the padding; ragged rows
    ("padded", 32, 71, 520),    # three 256-column wave spans, the last one partial
    ("basic", 1, 160, 317),     # unaligned rows: one node per thread
    ("basic", 1, 97, 516),      # aligned unpadded rows take the chunked path
    ("shared", 32, 160, 320)])
def test_step2d_bitexact_with_numpy_fd_step(variant, pad, NX, NY):
    from neuralmech_amd import reference_steps as rs

    dx, dy, dt, Mp, u = _setup(NX, NY, pad)
    a0, a1 = u.copy(), u.copy()
    g0 = torch.as_tensor(u).cuda()
    g1 = g0.clone()
    fn = {"padded": rs.padded_step, "shared": rs.shared_step}.get(variant)
    for _ in range(200):
        slice_ref.fd_step_reference_f32(a0, a1, a0, dt, dx, dy, 1.0, NX, NY)
        a0, a1 = a1, a0
        if variant == "basic":
            rs.basic_step(g0, g1, g0, 1.0, dt, dx, dy, NX, NY)
        else:
            fn(g0, g1, g0, 1.0, dt, dx, dy, NX, NY, Mp)
        g0, g1 = g1, g0
    torch.cuda.synchronize()
    assert np.array_equal(g1.cpu().numpy(), a1)
    assert np.array_equal(g0.cpu().numpy(), a0)


def test_step1d_bitexact():
    from neuralmech_amd import reference_steps as rs

    n = 1000
    x = np.linspace(0, 1, n)
    u = np.exp(-((x - 0.4) ** 2) / (2 * 0.03 ** 2)).astype(np.float32)
    dx = 1 / (n - 1)
    dt = 0.9 * dx
    a0, a1 = u.copy(), u.copy()
    g0 = torch.as_tensor(u).cuda()
    g1 = g0.clone()
    for _ in range(300):
        slice_ref.step1d_reference(a0, a1, a0, dt, dx, 1.0)
        a0, a1 = a1, a0
        rs.step1d(g0, g1, g0, 1.0, dt, dx, n)
        g0, g1 = g1, g0
    torch.cuda.synchronize()
    assert np.array_equal(g1.cpu().numpy(), a1)


@pytest.mark.parametrize("steps", [2, 4, 8])
def test_temporal_blocking_bit_identical_to_single_steps(steps):
    from neuralmech_amd import reference_steps as rs

    NX, NY = 450, 700
    dx, dy, dt, Mp, u = _setup(NX, NY)
    U = torch.as_tensor(np.stack([u, u])).cuda()
    V = torch.zeros_like(U)
    W = U.clone()
    u0, u1, v0, v1 = U[0], U[1], V[0], V[1]
    w0, w1 = W[0], W[1]
    total = 8 * steps
    for _ in range(total // steps):
        rs.temporal_step(u0, u1, v0, v1, 1.0, dt, dx, dy, NX, NY, Mp, steps)
        u0, u1, v0, v1 = v0, v1, u0, u1
    for _ in range(total):
        rs.padded_step(w0, w1, w0, 1.0, dt, dx, dy, NX, NY, Mp)
        w0, w1 = w1, w0
    torch.cuda.synchronize()
    assert float((u1 - w1).abs().max()) == 0.0
    assert torch.equal(u1, w1) and torch.equal(u0, w0)


def test_mixed_fp16_bitexact():
    from neuralmech_amd import reference_steps as rs

    NX, NY = 200, 398
    dx, dy, dt, Mp, u = _setup(NX, NY, 32)
    h = u.astype(np.float16)
    a0, a1 = h.copy(), h.copy()
    g0 = torch.as_tensor(h).cuda()
    g1 = g0.clone()
    for _ in range(150):
        slice_ref.mixed_step_reference(a0, a1, a0, dt, dx, dy, 1.0, NX, NY)
        a0, a1 = a1, a0
        rs.mixed_step(g0, g1, g0, 1.0, dt, dx, dy, NX, NY, Mp)
        g0, g1 = g1, g0
    torch.cuda.synchronize()
    assert np.array_equal(g1.cpu().numpy().view(np.uint16), a1.view(np.uint16))
    full = np.zeros_like(u)
    full[:NX, :NY] = slice_ref.gaussian_ic(NX, NY, center=(0.5, 0.5), sigma=0.05)
    a, b = full.copy(), full.copy()
    for _ in range(150):
        slice_ref.fd_step_reference_f32(a, b, a, dt, dx, dy, 1.0, NX, NY)
        a, b = b, a
    err = np.linalg.norm(a1.astype(np.float64) - b) / np.linalg.norm(b)
    assert err < 5e-2  # storage precision, not an arithmetic mismatch
    print(f"fp16 storage vs fp32 after 150 steps: relative L2 {err:.2e}")


def test_scaled_int16_bitexact():
    from neuralmech_amd import reference_steps as rs

    NX, NY = 130, 250
    dx, dy, dt, _, u = _setup(NX, NY, 128)
    Mp = u.shape[1]
    q, s = slice_ref.scaled_encode(u, NY)
    q[0], s[0], q[-1], s[-1] = 0, 0, 0, 0
    aq0, as0, aq1, as1 = q.copy(), s.copy(), q.copy(), s.copy()
    gq0, gs0 = torch.as_tensor(q).cuda(), torch.as_tensor(s).cuda()
    gq1, gs1 = gq0.clone(), gs0.clone()
    for _ in range(100):
        slice_ref.scaled_step_reference(aq0, as0, aq1, as1, aq0, as0, dt, dx, dy, 1.0, NX, NY)
        aq0, as0, aq1, as1 = aq1, as1, aq0, as0
        rs.scaled_step(gq0, gs0, gq1, gs1, gq0, gs0, 1.0, dt, dx, dy, NX, NY, Mp)
        gq0, gs0, gq1, gs1 = gq1, gs1, gq0, gs0
    torch.cuda.synchronize()
    assert np.array_equal(gq1.cpu().numpy(), aq1)
    assert np.array_equal(gs1.cpu().numpy(), as1)
    dec = slice_ref.scaled_decode(aq1, as1)
    assert np.isfinite(dec).all() and math.isfinite(float(np.abs(dec).max()))


def test_roofline_and_scaling_harness_small():
    import math

    from neuralmech_amd import harness

    r = harness.roofline(nx=66, ny=260, steps=80)
    assert r["bandwidth_GBs"] > 100 and r["peak_Gflops"] > 1000
    assert [k["kernel"] for k in r["kernels"]] == list(harness.KERNELS)
    assert all(math.isfinite(k["gflops"]) and k["gflops"] > 0 for k in r["kernels"])
    dofs, t = harness.scaling("padded", max_resolution=40, steps=40)
    assert dofs and len(dofs) == len(t) and all(x > 0 for x in t)


def test_stream_bench_kernels_compute_what_they_state():
    """wfd_stream_bench (the roofline harness's streaming ceilings): copy, a + b + c with and
    without the nontemporal hint, and the block sums of the read-only kernel; argument checks."""
    import ctypes

    import torch

    from neuralmech_amd import _lib

    lib = _lib.load()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    n = 4096 * 3 + 8  # a partial last block in every mode
    g = torch.Generator().manual_seed(3)
    a, b, c = (torch.rand(n, generator=g).cuda() for _ in range(3))
    o = torch.zeros(n, device="cuda")
    _lib.check(lib.wfd_stream_bench(a.data_ptr(), 0, 0, o.data_ptr(), n, 0, st))
    assert torch.equal(o, a)
    for mode in (1, 2):
        o.zero_()
        _lib.check(lib.wfd_stream_bench(a.data_ptr(), b.data_ptr(), c.data_ptr(), o.data_ptr(), n, mode, st))
        assert torch.equal(o, a + b + c)
    o.zero_()
    _lib.check(lib.wfd_stream_bench(a.data_ptr(), 0, 0, o.data_ptr(), n, 3, st))
    torch.cuda.synchronize()
    blocks = -(-n // 4096)
    ref = torch.stack([a[k * 4096:(k + 1) * 4096].double().sum() for k in range(blocks)])
    assert torch.allclose(o[:blocks].double(), ref, rtol=1e-5) and torch.all(o[blocks:] == 0)
    assert lib.wfd_stream_bench(a.data_ptr(), 0, 0, o.data_ptr(), n, 1, st) == _lib.WFD_EINVAL
    assert lib.wfd_stream_bench(a.data_ptr(), 0, 0, o.data_ptr(), n - 2, 0, st) == _lib.WFD_EINVAL
    assert lib.wfd_stream_bench(a.data_ptr() + 4, 0, 0, o.data_ptr(), 8, 0, st) == _lib.WFD_EINVAL
    assert lib.wfd_stream_bench(a.data_ptr(), 0, 0, o.data_ptr(), n, 4, st) == _lib.WFD_EINVAL


def test_storage_kernels_satisfy_the_format_pins():
    """One GPU step of the fp16 and int16 drop-ins checked against the formats' defining
    properties (tests/test_storage_pins.py): nearest binary16 of the fp32 update; |q| = 32767 at
    each block maximum, half-quantum decoding, nearest-integer q.  In rows 3..9 the int16 input is
    u1 = c, u0 = 2c (c = 1e-33, exactly cancelling in -u0 + 2 u1), so rows 4..8 update to
    coef * (-c / dy^2) = -7.4e-35 next to the boundary columns and to 0 inside: blocks whose maximum
    lies below 32767 / FLT_MAX, which encode as q = 0."""
    from neuralmech_amd import reference_steps as rs
    from tests.test_storage_pins import _int16_pins, _is_nearest_f16

    NX, NY, Mp = 40, 120, 128
    dt, dx, dy = 0.3, 1.0, 1.1
    rng = np.random.default_rng(9)
    u1 = np.zeros((NX, Mp), np.float32)
    u1[18:-1, 1:NY - 1] = rng.standard_normal((NX - 19, NY - 2))
    u0 = 0.9 * u1
    # fp16
    h0, h1 = u0.astype(np.float16), u1.astype(np.float16)
    g0, g1 = torch.as_tensor(h0).cuda(), torch.as_tensor(h1).cuda()
    rs.mixed_step(g0, g1, g0, 1.0, dt, dx, dy, NX, NY, Mp)
    a0, a1 = h0[:, :NY].astype(np.float32), h1[:, :NY].astype(np.float32)
    x = a0.copy()
    slice_ref.fd_step_reference(a0, a1, x, dt, dx, dy, 1.0)
    out = g0.cpu().numpy()
    assert _is_nearest_f16(out[1:NX - 1, 1:NY - 1], x[1:-1, 1:-1]).all()
    # int16
    c = np.float32(1e-33)
    u1[3:10, 1:NY - 1] = c
    u0[3:10, 1:NY - 1] = 2 * c
    q0, s0 = slice_ref.scaled_encode(u0, Mp)
    q1, s1 = slice_ref.scaled_encode(u1, Mp)
    gq0, gs0 = torch.as_tensor(q0).cuda(), torch.as_tensor(s0).cuda()
    gq1, gs1 = torch.as_tensor(q1).cuda(), torch.as_tensor(s1).cuda()
    rs.scaled_step(gq0, gs0, gq1, gs1, gq0, gs0, 1.0, dt, dx, dy, NX, NY, Mp)
    torch.cuda.synchronize()
    q2, s2 = gq0.cpu().numpy(), gs0.cpu().numpy()
    d0, d1 = slice_ref.scaled_decode(q0, s0), slice_ref.scaled_decode(q1, s1)
    x = np.zeros_like(d1)
    slice_ref.fd_step_reference(d0[:, :NY], d1[:, :NY], x[:, :NY], dt, dx, dy, 1.0)
    x[:, 0] = 0
    assert all(_int16_pins(q2[18:-1], s2[18:-1], x[18:-1]))
    blk = np.abs(x[4:9]).reshape(5, 4, 32).max(axis=2)
    limit = 32767 / np.finfo(np.float32).max
    assert np.all((blk[:, [0, 3]] > 0) & (blk[:, [0, 3]] < limit)) and np.all(blk[:, [1, 2]] == 0)
    assert np.all(q2[4:9] == 0) and np.all(np.isfinite(s2[4:9]))
