"""CPU pins of the storage-compressed restatements (fp16 mixed_step, int16 block-scaled
scaled_step of projects/18_simple_wave_fd/step2D_wave.cu:85-183).  Each property is checked from
its definition, not by re-running the restatement's formula, and a planted mistake in a mutated
copy of the restatement (on the CPU) must fail it:

* fp16: every stored value is the binary16 nearest to the fp32 update (ties to an even
  significand); a truncating store fails.
* int16: per 32-node block, |q| = 32767 at the block's largest magnitude, every value decodes
  within half a quantum (m / 32767 / 2) of the fp32 value it encodes, q is the nearest integer
  (ties to even); 31-node blocks and truncation fail.
* Both against fp32 at the reference's own self-check times (2D_wave_cuda_scaled.py:121-126):
  relative L2 errors below the bounds the two formats' rounding allows.
"""

import numpy as np
import pytest

from oracle import slice_ref


def _fields(N=40, M=70, Mp=128, seed=0):
    rng = np.random.default_rng(seed)
    x = np.linspace(0, 1, M)[None, :]
    y = np.linspace(0, 1, N)[:, None]
    base = np.sin(3 * np.pi * x) * np.sin(2 * np.pi * y) + 0.2 * rng.standard_normal((N, M))
    u0 = np.zeros((N, Mp), np.float32)
    u1 = np.zeros((N, Mp), np.float32)
    u0[1:-1, 1:M - 1] = base[1:-1, 1:-1]
    u1[1:-1, 1:M - 1] = 1.01 * base[1:-1, 1:-1]
    return u0, u1, N, M, Mp


# ------------------------------------------------------------------------------ fp16
def _is_nearest_f16(h: np.ndarray, x: np.ndarray) -> np.ndarray:
    """True where the binary16 value h is a nearest binary16 to the fp32 value x, ties resolved to
    the even significand (the neighbours are found by stepping the bit pattern, not by a cast)."""
    h = h.astype(np.float16)
    with np.errstate(over="ignore"):
        up = np.nextafter(h, np.float16(np.inf))
        dn = np.nextafter(h, np.float16(-np.inf))
    xd = x.astype(np.float64)
    d = np.abs(h.astype(np.float64) - xd)
    du = np.abs(up.astype(np.float64) - xd)
    ddn = np.abs(dn.astype(np.float64) - xd)
    nearest = (d <= du) & (d <= ddn)
    tie = (d == du) | (d == ddn)
    even = (h.view(np.uint16) & 1) == 0
    return nearest & (~tie | even)


def _mixed_truncating(u0, u1, u2, dt, dx, dy, c, N, M):
    """A planted mistake: the mixed step storing with rounding toward zero."""
    a0 = u0[:N, :M].astype(np.float32)
    a1 = u1[:N, :M].astype(np.float32)
    a2 = a0.copy()
    slice_ref.fd_step_reference(a0, a1, a2, dt, dx, dy, c)
    v = a2[1:-1, 1:-1]
    h = v.astype(np.float16)
    over = np.abs(h.astype(np.float32)) > np.abs(v)  # rounded away from zero: step back
    h[over] = np.nextafter(h[over], np.float16(0))
    u2[1:N - 1, 1:M - 1] = h
    return u2


@pytest.mark.parametrize("mutant", [False, True])
def test_fp16_store_is_the_nearest_binary16(mutant):
    u0, u1, N, M, Mp = _fields()
    h0, h1 = u0.astype(np.float16), u1.astype(np.float16)
    dt, dx, dy = 0.3, 1.0, 1.1
    # the fp32 update from the decoded inputs (decoding binary16 -> fp32 is exact)
    a0, a1 = h0[:, :M].astype(np.float32), h1[:, :M].astype(np.float32)
    assert np.array_equal(a0.astype(np.float16), h0[:, :M])
    x = a0.copy()
    slice_ref.fd_step_reference(a0, a1, x, dt, dx, dy, 1.0)
    out = h0.copy()
    step = _mixed_truncating if mutant else slice_ref.mixed_step_reference
    step(h0, h1, out, dt, dx, dy, 1.0, N, M)
    ok = _is_nearest_f16(out[1:N - 1, 1:M - 1], x[1:-1, 1:-1])
    if mutant:
        assert not ok.all(), "a truncating store passed the nearest-binary16 pin"
    else:
        assert ok.all()
        # the boundary ring and the row padding are never written
        assert np.array_equal(out[[0, N - 1]], h0[[0, N - 1]]) and np.array_equal(out[:, M - 1:], h0[:, M - 1:])
        assert np.array_equal(out[:, 0], h0[:, 0])


def test_nearest_pin_decides_ties_and_overflow():
    """Exact binary16 midpoints go to the even significand; values past the largest finite
    binary16 round to infinity only beyond the midpoint (65520)."""
    one = np.float16(1.0)
    nxt = np.nextafter(one, np.float16(2))
    mid = np.float32((np.float64(one) + np.float64(nxt)) / 2)
    assert _is_nearest_f16(np.array([one]), np.array([mid])).all()
    assert not _is_nearest_f16(np.array([nxt]), np.array([mid])).all()  # odd significand
    assert _is_nearest_f16(np.array([np.float16(65504)]), np.array([np.float32(65519)])).all()


# ----------------------------------------------------------------------------- int16
def _encode_blocks(u, width, rounding=np.rint):
    """Block-scaled encoding with a given block width and rounding (for planting mistakes)."""
    v = u.astype(np.float32)
    nb = v.shape[1] // width
    blk = np.abs(v[:, : nb * width]).reshape(v.shape[0], nb, width).max(axis=2)
    with np.errstate(divide="ignore", over="ignore"):
        inv = (np.float32(32767.0) / blk).astype(np.float32)
    inv[~np.isfinite(inv)] = 0
    q = np.zeros(v.shape, np.int16)
    q[:, : nb * width] = rounding(v[:, : nb * width] * np.repeat(inv, width, axis=1)).astype(np.int16)
    return q, (blk / np.float32(32767.0)).astype(np.float32)


def _int16_pins(q, s, u):
    """The block-scaled format's defining properties for values u (fp32) encoded as (q, s)."""
    nb = s.shape[1]
    uq = u[:, : nb * 32].reshape(u.shape[0], nb, 32)
    qq = q[:, : nb * 32].reshape(q.shape[0], nb, 32).astype(np.int64)
    m = np.abs(uq).max(axis=2)
    nz = m > 0
    # |q| = 32767 at the block's largest magnitude, and nowhere above it
    at_max = np.take_along_axis(np.abs(qq), np.abs(uq).argmax(axis=2)[..., None], 2)[..., 0]
    p1 = np.all(at_max[nz] == 32767) and np.abs(qq).max() <= 32767
    # every value within half a quantum of its decoded value (fp32 decode: one more rounding)
    dec = np.repeat(s, 32, axis=1)[:, : nb * 32].reshape(uq.shape) * qq.astype(np.float32)
    quantum = (m / 32767.0)[..., None]
    err = np.abs(dec.astype(np.float64) - uq.astype(np.float64))
    p2 = np.all(err <= 0.5 * quantum * (1 + 1e-6) + 4 * np.finfo(np.float32).eps * np.abs(uq))
    # q is the nearest integer to u * (32767 / m) in fp32, ties to even
    inv = np.where(nz, np.float32(32767.0) / np.where(nz, m, 1).astype(np.float32), 0).astype(np.float32)
    t = (uq * inv[..., None]).astype(np.float64)
    d = np.abs(qq - t)
    p3 = np.all(d <= 0.5) and np.all((d != 0.5) | (qq % 2 == 0))
    return bool(p1), bool(p2), bool(p3)


@pytest.mark.parametrize("mutant", [None, "block31", "trunc"])
def test_int16_encoding_pins(mutant):
    rng = np.random.default_rng(3)
    u = (rng.standard_normal((12, 128)) * np.exp(rng.uniform(-30, 10, (12, 1)))).astype(np.float32)
    u[3, 32:64] = 0.0  # an all-zero block encodes as q = 0, s = 0
    if mutant is None:
        q, s = slice_ref.scaled_encode(u, 128)
        assert np.all(q[3, 32:64] == 0) and s[3, 1] == 0
    elif mutant == "block31":
        q, s = _encode_blocks(u, 31)  # read as the 32-node format it claims to be
        s = s[:, : u.shape[1] // 32]
    else:
        q, s = _encode_blocks(u, 32, np.trunc)
    p = _int16_pins(q, s, u)
    if mutant is None:
        assert all(p), p
    else:
        assert not all(p), f"planted mistake {mutant} passed the int16 pins {p}"


def test_int16_rounding_ties_go_to_even():
    u = np.zeros((1, 32), np.float32)
    u[0, 0] = 32767.0
    u[0, 1] = 2.5
    u[0, 2] = 3.5
    u[0, 3] = -2.5
    q, s = slice_ref.scaled_encode(u, 32)
    assert s[0, 0] == 1.0 and list(q[0, :4]) == [32767, 2, 4, -2]


def test_scaled_step_stores_satisfy_the_pins():
    u0, u1, N, M, Mp = _fields(seed=7)
    dt, dx, dy = 0.3, 1.0, 1.1
    q0, s0 = slice_ref.scaled_encode(u0, Mp)
    q1, s1 = slice_ref.scaled_encode(u1, Mp)
    d0, d1 = slice_ref.scaled_decode(q0, s0), slice_ref.scaled_decode(q1, s1)
    x = np.zeros_like(d1)
    slice_ref.fd_step_reference(d0[:N, :M], d1[:N, :M], x[:N, :M], dt, dx, dy, 1.0)
    q2, s2 = q0.copy(), s0.copy()
    slice_ref.scaled_step_reference(q0, s0, q1, s1, q2, s2, dt, dx, dy, 1.0, N, M)
    assert all(_int16_pins(q2[1:-1], s2[1:-1], x[1:-1]))
    # rows 0 and N-1 untouched; boundary columns and padding of interior rows encode zero
    assert np.array_equal(q2[[0, N - 1]], q0[[0, N - 1]])
    assert np.all(q2[1:-1, 0] == 0) and np.all(q2[1:-1, M - 1:] == 0)


def test_tiny_block_maximum_encodes_zero():
    """For 0 < m < 32767 / FLT_MAX (about 9.6e-35) 32767 / m overflows fp32.  The reference
    (step2D_wave.cu:179) then multiplies by inf and converts with saturation, storing the low 16
    bits of the saturated integer; this restatement (and the kernel) store q = 0 for the block, the
    value it decodes closest to.  Above that range both agree."""
    u = np.zeros((1, 32), np.float32)
    u[0, 5] = 5e-35
    q, s = slice_ref.scaled_encode(u, 32)
    assert np.all(q == 0) and np.isfinite(s).all()
    u[0, 5] = 1e-33
    q, s = slice_ref.scaled_encode(u, 32)
    assert q[0, 5] == 32767 and abs(float(s[0, 0]) * 32767 / 1e-33 - 1) < 1e-6


@pytest.mark.parametrize("kind", ["fp16", "int16"])
def test_error_against_fp32_at_the_reference_self_check_times(kind):
    """The reference compares its fp16 and int16 runs with fp32 at t = 1 and t = 3
    (2D_wave_cuda_scaled.py:121-126).  Here, on a 96 x 128 Gaussian pulse with c dt / h = 0.5,
    the relative L2 error after as many steps stays below the per-step storage rounding (2^-11
    for fp16, 0.5 / 32767 of the block maximum for int16) times the step count."""
    N, M, Mp = 96, 128, 128
    y, x = np.mgrid[0:N, 0:M]
    u = np.exp(-((x - 64) ** 2 + (y - 48) ** 2) / 40.0).astype(np.float32)
    u[[0, -1]] = 0
    u[:, [0, M - 1]] = 0
    dx = dy = 1.0
    dt = 0.5
    steps = [int(1.0 / dt * 10), int(3.0 / dt * 10)]  # t = 1, 3 in units of 10 steps per second
    f0, f1 = u.copy(), u.copy()
    if kind == "fp16":
        a0, a1 = u.astype(np.float16), u.astype(np.float16)
        bound = 2.0 ** -11
    else:
        (a0, b0), (a1, b1) = slice_ref.scaled_encode(u, Mp), slice_ref.scaled_encode(u, Mp)
        bound = 0.5 / 32767 * 8  # half a quantum of the block max, against a field ~1/8 of it in L2
    done = 0
    for n in steps:
        for _ in range(n - done):
            slice_ref.fd_step_reference(f0, f1, f0, dt, dx, dy, 1.0)
            f0, f1 = f1, f0
            if kind == "fp16":
                slice_ref.mixed_step_reference(a0, a1, a0, dt, dx, dy, 1.0, N, M)
                a0, a1 = a1, a0
            else:
                slice_ref.scaled_step_reference(a0, b0, a1, b1, a0, b0, dt, dx, dy, 1.0, N, M)
                (a0, b0), (a1, b1) = (a1, b1), (a0, b0)
        done = n
        got = a1.astype(np.float32) if kind == "fp16" else slice_ref.scaled_decode(a1, b1)
        err = np.linalg.norm(got - f1) / np.linalg.norm(f1)
        assert 0 < err < bound * n, (kind, n, err)
