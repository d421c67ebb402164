"""CPU checks of the uneven batch's host side: the per-shot op list that the persistent kernel
runs, replayed symbolically (which time level, or which level's Laplacian, every field holds),
and argument validation."""

import ctypes

import numpy as np
import pytest

from neuralmech_amd import _lib
from neuralmech_amd.batch import shot_ops

K_GATHER, K_STEP, K_ADJ = 0, 1, 2
F_INJECT, F_FLUSH, F_FIRST, F_GREC, F_GSRC = 1, 2, 4, 8, 16
ZERO = ("zero",)


def _holds(tag, kind, k, N):
    """Does a field tagged `tag` hold level k of u (kind 'u') or mu (kind 'mu')?"""
    if tag == (kind, k):
        return True
    # the store starts zeroed: u^{-1} = u^0 = 0 and mu^{N+1} = mu^{N+2} = 0
    zero_level = k <= 0 if kind == "u" else k >= N + 1
    return tag == ZERO and zero_level


def replay(N, seg, has_src, has_rec):
    ops, nf = shot_ops(N, seg, has_src, has_rec)
    assert np.all(np.diff(ops[:, 10]) >= 0) and ops[0, 10] == 0  # ops in phase order
    tag = [ZERO] * nf
    rec_rows, src_rows, flushes, firsts = [], [], [], []
    adj_seen = []
    fwd_done = False
    gpart = None
    for ph in range(int(ops[-1, 10]) + 1):
        group = ops[ops[:, 10] == ph].tolist()
        assert 1 <= len(group) <= 2
        # ops of one phase run concurrently: no op writes what another op of the phase touches
        writes = [w for o in group if o[0] != K_GATHER for w in ((o[4], o[5]) if o[0] == K_STEP else (o[4],)) if w >= 0]
        touched = [[f for f in (o[2], o[3], o[4], o[5], o[7]) if f >= 0] for o in group]
        for i, o in enumerate(group):
            own = [w for w in ((o[4], o[5]) if o[0] == K_STEP else (o[4],)) if w >= 0] if o[0] != K_GATHER else []
            for j, t in enumerate(touched):
                if j != i:
                    assert not set(own) & set(t), (ph, group)
        assert len(writes) == len(set(writes))
        pending = []
        for kind, flags, up, uc, un, uf, gp, gf, step, row, _ in group:
            # gather and stencil of one op run concurrently: the gathered field is not written
            assert not (flags & (F_GREC | F_GSRC)) or kind == K_GATHER or gf != un
            if flags & F_GREC:
                assert _holds(tag[gf], "u", row + 1, N), (row, tag[gf])
                rec_rows.append(row)
            if flags & F_GSRC:
                assert _holds(tag[gf], "mu", row + 1, N), (row, tag[gf])
                src_rows.append(row)
            if kind == K_STEP:
                n = step
                assert un != uc and 0 <= un < nf
                assert _holds(tag[up], "u", n - 1, N) and _holds(tag[uc], "u", n, N), (n, tag[up], tag[uc])
                assert bool(flags & F_INJECT) == has_src
                pending.append((un, ("u", n + 1)))
                if uf >= 0:  # also stores L u^n, the Laplacian of its current level
                    assert uf not in (up, uc, un)
                    pending.append((uf, ("Lu", n)))
                if n == N - 1:
                    fwd_done = True
            elif kind == K_ADJ:
                m = step + 1
                assert fwd_done and un != uc
                assert _holds(tag[up], "mu", m + 2, N) and _holds(tag[uc], "mu", m + 1, N), (m, tag[up], tag[uc])
                assert tag[uf] == ("Lu", m - 1), (m, tag[uf])
                assert bool(flags & F_INJECT) == has_rec
                assert not adj_seen or m == adj_seen[-1] - 1
                assert gpart is None or gp == gpart
                gpart = gp
                adj_seen.append(m)
                pending.append((un, ("mu", m)))
                if flags & F_FLUSH:
                    flushes.append(m - 1)
                if flags & F_FIRST:
                    firsts.append(m - 1)
            else:
                assert kind == K_GATHER and flags & (F_GREC | F_GSRC)
        for f, t in pending:
            tag[f] = t
    # every adjoint step once, newest first; flushes where the gradient plan flushes
    assert adj_seen == list(range(N, 0, -1))
    assert flushes == [m - 1 for m in range(N, 0, -1) if (m - 1) % 32 == 0]
    assert firsts == flushes[:1]
    assert sorted(rec_rows) == (list(range(N)) if has_rec else [])
    assert sorted(src_rows) == (list(range(N)) if has_src else [])
    # the imaging partial never doubles as a wave field
    for kind, flags, up, uc, un, uf, gp, gf, step, row, _ in ops.tolist():
        assert gpart is None or gpart not in (up, uc, un, uf, gf)
    return ops, nf


@pytest.mark.parametrize("N,seg", [(1, 0), (2, 0), (2, 1), (3, 1), (7, 0), (7, 3), (7, 7), (33, 0), (64, 5),
                                   (65, 8), (100, 1), (100, 100), (257, 0), (500, 31)])
def test_schedule_replays(N, seg):
    replay(N, seg, True, True)


@pytest.mark.parametrize("has_src,has_rec", [(False, True), (True, False), (False, False)])
def test_schedule_without_points(has_src, has_rec):
    replay(45, 6, has_src, has_rec)


@pytest.mark.parametrize("N,seg", [(10, 1), (10, 3), (3000, 0), (5000, 0)])
def test_schedule_length_and_fields(N, seg):
    ops, nf = shot_ops(N, seg)

    def fields(C):
        nck = -(-N // C)
        return (nck + 1 if C == 1 else 2 * nck) + (0 if C == 1 else (2 if nck >= 2 else 1) * C) + 5

    C = seg or 1 + int(np.argmin([fields(c) for c in range(1, N + 1)]))
    nck = -(-N // C)
    last = N - (nck - 1) * C  # steps of the last segment
    # ops: forward + lone gather (C = 1: no recompute to merge it into) + recompute (every level
    # once, storing its Laplacian) + adjoint + final gather
    assert len(ops) == N + (1 if C == 1 else 0) + (0 if C == 1 else N) + N + 1
    assert nf == fields(C)
    # phases: forward, the last segment's recompute alone, then each adjoint segment together with
    # the previous segment's recompute, then the final gather
    rec = C if C > 1 else 0
    phases = N + (last if C > 1 else 1) + sum(max(min(C, N - k * C), rec if k else 0) for k in range(nck)) + 1
    assert ops[-1, 10] + 1 == phases


def test_shot_ops_rejects_bad_arguments():
    lib = _lib.load()
    cnt, nf = ctypes.c_int64(), ctypes.c_int32()
    assert lib.wfd_batch_shot_ops(0, 0, 1, 1, None, 0, ctypes.byref(cnt), ctypes.byref(nf)) == _lib.WFD_EINVAL
    assert lib.wfd_batch_shot_ops(10, 11, 1, 1, None, 0, ctypes.byref(cnt), ctypes.byref(nf)) == _lib.WFD_EINVAL
    out = np.zeros((2, 11), np.int32)
    rc = lib.wfd_batch_shot_ops(10, 0, 1, 1, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), 2,
                                ctypes.byref(cnt), ctypes.byref(nf))
    assert rc == _lib.WFD_EINVAL and "needed" in _lib.last_error()


def test_batch_create_validates_without_gpu():
    lib = _lib.load()
    arr = (_lib.BatchShot * 1)()
    n = (ctypes.c_int64 * 3)(64, 64, 64)
    h = (ctypes.c_double * 3)(10.0, 10.0, 1.0)
    assert lib.wfd_grid_init(ctypes.byref(arr[0].grid), 3, n, h, 1e-3, 8, _lib.WFD_F32, 1) == 0
    arr[0].nsteps = 10
    b = ctypes.c_void_p()
    assert lib.wfd_batch_create(ctypes.byref(b), 1, arr, None) == _lib.WFD_EINVAL
    assert "2D" in _lib.last_error()
    assert lib.wfd_grid_init(ctypes.byref(arr[0].grid), 2, n, h, 1e-3, 8, _lib.WFD_F32, 1) == 0
    assert lib.wfd_batch_create(ctypes.byref(b), 1, arr, None) == _lib.WFD_EINVAL
    assert "coef" in _lib.last_error()
    assert lib.wfd_batch_create(ctypes.byref(b), 0, arr, None) == _lib.WFD_EINVAL
