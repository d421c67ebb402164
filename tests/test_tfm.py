"""Total focusing method: the analytic signal against scipy.signal.hilbert, the CPU restatement
(oracle/tfm_ref.py) pinned by a point scatterer's synthetic capture, an explicit per-pixel loop
and the agreement of the two forms on reciprocal traces; the HIP imager against the restatement
(GPU)."""

import math

import numpy as np
import pytest
import torch

from oracle import tfm_ref


def _capture(E=16, Nt=400, c=1.0, dt=0.004, t0=0.05, scatterer=(0.47, 0.38), f=12.0, seed=0):
    """Analytic traces of a point scatterer at `scatterer` seen by E elements on y = 1: a Gabor pulse
    delayed by the two-way time of flight (reciprocal: h_er = h_re)."""
    xe = np.linspace(0.1, 0.9, E)
    el = np.stack([xe, np.ones(E)], 1)
    d = np.hypot(xe - scatterer[0], 1.0 - scatterer[1]) / c
    t = np.arange(Nt) * dt
    tof = t0 + d[:, None, None] + d[None, :, None]
    s = t[None, None, :] - tof
    env = np.exp(-0.5 * (s * f * 2.0) ** 2)
    h = env * np.exp(2j * math.pi * f * s)
    return h, el, t0, dt, c


def test_analytic_signal_is_scipy_hilbert():
    from neuralmech_amd.tfm import analytic_signal

    rng = np.random.default_rng(0)
    for n in (128, 129):
        x = rng.standard_normal((3, n, 5))
        got = analytic_signal(torch.as_tensor(x), axis=1).numpy()
        ref = tfm_ref.hilbert(x, axis=1)
        assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
        assert np.allclose(got.real, x)


@pytest.mark.parametrize("form", ["all", "reciprocal"])
def test_point_scatterer_focuses(form):
    h, el, t0, dt, c = _capture()
    x, y = np.meshgrid(np.linspace(0, 1, 101), np.linspace(0, 1, 101), indexing="ij")
    f = tfm_ref.tfm_all_pairs if form == "all" else tfm_ref.tfm_reciprocal
    img = f(h, el, t0, dt, c, x, y)
    i, j = np.unravel_index(np.argmax(img), img.shape)
    assert abs(x[i, j] - 0.47) <= 0.011 and abs(y[i, j] - 0.38) <= 0.011
    # at the scatterer every pair adds in phase: E^2 times the pulse peak
    assert img[i, j] > 0.9 * h.shape[0] ** 2


def test_all_pairs_equals_explicit_loop():
    h, el, t0, dt, c = _capture(E=6, Nt=300)
    pts = [(0.2, 0.3), (0.47, 0.38), (0.9, 0.95)]
    x = np.array([p[0] for p in pts])
    y = np.array([p[1] for p in pts])
    img = tfm_ref.tfm_all_pairs(h, el, t0, dt, c, x, y)
    E, _, Nt = h.shape
    for q, (px, py) in enumerate(pts):
        acc = 0j
        for e in range(E):
            for r in range(E):
                t = (t0 + math.hypot(px - el[e, 0], py - el[e, 1]) / c + math.hypot(px - el[r, 0], py - el[r, 1]) / c) / dt
                k = int(t)
                if k >= Nt - 2:
                    continue
                w = t - k
                acc += (1 - w) * h[e, r, k] + w * h[e, r, k + 1]
        assert abs(abs(acc) - img[q]) <= 1e-12 * max(1.0, abs(acc))


def test_forms_agree_on_reciprocal_traces():
    h, el, t0, dt, c = _capture(E=8, Nt=500)
    x, y = np.meshgrid(np.linspace(0, 1, 33), np.linspace(0.2, 0.9, 29), indexing="ij")
    a = tfm_ref.tfm_all_pairs(h, el, t0, dt, c, x, y)
    b = tfm_ref.tfm_reciprocal(h, el, t0, dt, c, x, y)
    assert np.abs(a - b).max() < 1e-10 * a.max()


@pytest.mark.gpu
@pytest.mark.parametrize("E", [8, 32, 40])
@pytest.mark.parametrize("form", ["all", "reciprocal"])
def test_gpu_tfm_matches_restatement(E, form):
    from neuralmech_amd import tfm

    h, el, t0, dt, c = _capture(E=E, Nt=420)
    x, y = np.meshgrid(np.linspace(0, 1, 67), np.linspace(0, 1, 53), indexing="ij")
    if form == "all":
        # a non-reciprocal capture, so that the two forms differ
        h = h * (1.0 + 0.1 * np.arange(E)[:, None, None] / E)
        got = tfm.tfm(h, el, t0, dt, c, x, y)
        ref = tfm_ref.tfm_all_pairs(h.astype(np.complex64), el.astype(np.float32), np.float32(t0), np.float32(dt),
                                    np.float32(c), x.astype(np.float32), y.astype(np.float32))
    else:
        got = tfm.tfm_kernel(h, el, t0, dt, c, x, y)
        ref = tfm_ref.tfm_reciprocal(h.astype(np.complex64), el.astype(np.float32), np.float32(t0), np.float32(dt),
                                     np.float32(c), x.astype(np.float32), y.astype(np.float32))
    got = got.cpu().numpy()
    assert got.shape == x.shape
    err = np.abs(got - ref).max() / ref.max()
    print(f"TFM {form} E={E}: largest difference {err:.1e} of the image maximum")
    assert err < 2e-5


@pytest.mark.gpu
@pytest.mark.parametrize("side", [64, 200, 520, 1030])
def test_gpu_tfm_lane_groups(side):
    """Images from 4 K to 1 M pixels (16, 16, 4 and 1 lanes per pixel): both forms against the
    restatement.  The record (500 samples) holds every pulse: where a trace is still nonzero at
    the end of the record, the forms' end-of-record skip makes pixels whose delay rounds across
    the last sample differ between fp32 and the restatement's float64 delays."""
    from neuralmech_amd import tfm

    h, el, t0, dt, c = _capture(E=12, Nt=500)
    x, y = np.meshgrid(np.linspace(0, 1, side), np.linspace(0, 1, side + 3), indexing="ij")
    args32 = (h.astype(np.complex64), el.astype(np.float32), np.float32(t0), np.float32(dt), np.float32(c),
              x.astype(np.float32), y.astype(np.float32))
    for got, ref in ((tfm.tfm_kernel(h, el, t0, dt, c, x, y), tfm_ref.tfm_reciprocal(*args32)),
                     (tfm.tfm(h, el, t0, dt, c, x, y), tfm_ref.tfm_all_pairs(*args32))):
        got = got.cpu().numpy()
        assert got.shape == x.shape
        assert np.abs(got - ref).max() / ref.max() < 2e-5


@pytest.mark.gpu
def test_gpu_analytic_signal_fp32():
    from neuralmech_amd.tfm import analytic_signal

    x = np.random.default_rng(1).standard_normal((4, 6, 1001)).astype(np.float32)
    got = analytic_signal(torch.as_tensor(x).cuda(), axis=-1).cpu().numpy()
    ref = tfm_ref.hilbert(x.astype(np.float64), axis=-1)
    assert np.abs(got - ref).max() < 1e-5 * np.abs(ref).max()


def _end_of_record_bound(h, el, t0, dt, c, x, y, form, eps=2e-3):
    """Per pixel, the largest difference an fp32 evaluation may show against the float64 delays of
    the restatement at the end of the record: a pair whose delay t (in samples) lies within eps of
    the skip threshold (all pairs: t < Nt - 2 kept; reciprocal: t < Nt - 1 kept) may be kept by one
    and skipped by the other, which changes the sum by at most the pair's weight times the largest
    |h| of its samples next to the threshold.  Zero on pixels with no such pair."""
    E, _, Nt = h.shape
    T = Nt - 2 if form == "all" else Nt - 1
    tau = [np.hypot(x.astype(np.float64) - ex, y.astype(np.float64) - ey) / c for ex, ey in el.astype(np.float64)]
    bound = np.zeros(x.shape)
    for e in range(E):
        for r in range(E) if form == "all" else range(e, E):
            t = (t0 + tau[e] + tau[r]) / dt
            near = np.abs(t - T) < eps
            if near.any():
                wgt = 1.0 if form == "all" or r == e else 2.0
                bound += near * wgt * np.abs(h[e, r, T - 1:T + 1]).max()
    return bound


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["all", "reciprocal"])
def test_gpu_tfm_end_of_record(form):
    """A record cut while the pulses are still arriving (330 samples: the traces are large at the
    last samples), on a 1030^2 image: pixels where no pair's delay lies near the skip threshold
    match the restatement within 2e-5 of the image maximum; on the pixels where one does, the
    difference stays within the skipped pairs' contribution (``_end_of_record_bound``)."""
    from neuralmech_amd import tfm

    h, el, t0, dt, c = _capture(E=12, Nt=330)
    assert np.abs(h[:, :, -3:]).max() > 0.5  # the regime: nonzero at the end of the record
    x, y = np.meshgrid(np.linspace(0, 1, 1030), np.linspace(0, 1, 1033), indexing="ij")
    args32 = (h.astype(np.complex64), el.astype(np.float32), np.float32(t0), np.float32(dt), np.float32(c),
              x.astype(np.float32), y.astype(np.float32))
    if form == "all":
        got, ref = tfm.tfm(h, el, t0, dt, c, x, y), tfm_ref.tfm_all_pairs(*args32)
    else:
        got, ref = tfm.tfm_kernel(h, el, t0, dt, c, x, y), tfm_ref.tfm_reciprocal(*args32)
    got = got.cpu().numpy()
    bound = _end_of_record_bound(*args32, form)
    diff = np.abs(got - ref)
    tol = 2e-5 * ref.max()
    border = bound > 0
    print(f"\n[TFM end of record, {form}] {border.sum()} pixels near the threshold; elsewhere largest "
          f"difference {diff[~border].max() / ref.max():.1e} of the maximum; near it "
          f"{diff[border].max() / ref.max():.1e} (bound {bound[border].max() / ref.max():.1e})")
    assert border.sum() > 0
    assert diff[~border].max() <= tol
    assert np.all(diff[border] <= bound[border] + tol)
