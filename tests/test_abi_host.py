"""CPU tests of the C ABI: the library loads and exports every declared symbol; host-only
functions (layout, weights, stability limit, point lists, argument validation) behave; the
host-side layout helpers round-trip.  No kernel is launched here."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest
import torch

from neuralmech_amd import _lib, solver
from oracle import c_oracle, slice_ref

ROOT = Path(__file__).resolve().parent.parent


def _declared():
    text = (ROOT / "include" / "wavefd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wfd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 18
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} not bound in _lib.SIGNATURES"
    assert b"gfx950" in lib.wfd_version()


def _grid(shape, h, order=8, dtype=_lib.WFD_F32, nshots=1, dt=0.1):
    g = _lib.Grid()
    n = (ctypes.c_int64 * 3)(*shape, *([1] * (3 - len(shape))))
    hh = (ctypes.c_double * 3)(*h, *([1.0] * (3 - len(h))))
    rc = _lib.load().wfd_grid_init(ctypes.byref(g), len(shape), n, hh, dt, order, dtype, nshots)
    return rc, g


@pytest.mark.parametrize("shape", [(2048, 2048), (1024, 1024), (513, 77), (512, 512, 512), (21, 19, 35), (101,)])
def test_layout(shape):
    rc, g = _grid(shape, [1.0] * len(shape))
    assert rc == 0
    nf = shape[-1]
    assert g.ld % 32 == 0 and g.ld >= nf + 4
    if len(shape) == 3:
        assert g.plane >= shape[1] + 4
        assert g.offset == 4 * g.plane * g.ld and g.elems == (shape[0] + 8) * g.plane * g.ld
    else:
        rows = shape[0] if len(shape) == 2 else 1
        assert g.offset == 4 * g.ld and g.elems == (rows + 8) * g.ld


@pytest.mark.parametrize("bad", [dict(order=3), dict(order=10), dict(shape=(2, 10)), dict(dt=0.0),
                                 dict(dtype=7), dict(nshots=0)])
def test_grid_init_rejects(bad):
    kw = dict(shape=(10, 10), h=(1.0, 1.0), order=4, dtype=_lib.WFD_F32, nshots=1, dt=0.1)
    kw.update(bad)
    rc, _ = _grid(kw["shape"], kw["h"], kw["order"], kw["dtype"], kw["nshots"], kw["dt"])
    assert rc == _lib.WFD_EINVAL
    assert _lib.last_error()


@pytest.mark.parametrize("ndim,order", [(1, 2), (2, 4), (2, 8), (3, 6), (3, 8)])
@pytest.mark.parametrize("dt_code", [_lib.WFD_F32, _lib.WFD_F64])
def test_stencil_weights_match_oracle(ndim, order, dt_code):
    h = [0.7, 1.3, 2.1][:ndim]
    rc, g = _grid([9] * ndim, h, order, dt_code)
    assert rc == 0
    wc = ctypes.c_double()
    w = (ctypes.c_double * 12)()
    assert _lib.load().wfd_stencil_weights(ctypes.byref(g), ctypes.byref(wc), w) == 0
    npdt = np.float32 if dt_code == _lib.WFD_F32 else np.float64
    owc, ow = c_oracle.weights(ndim, h, order, npdt)
    assert wc.value == float(owc)
    assert np.array_equal(np.array(w[: ndim * (order // 2)]), ow.reshape(-1).astype(np.float64))


@pytest.mark.parametrize("ndim,order", [(2, 2), (2, 8), (3, 8), (1, 4)])
def test_stable_dt_matches_closed_form(ndim, order):
    h = [10.0, 12.0, 9.0][:ndim]
    assert solver.stable_dt(h, 4500.0, order) == pytest.approx(slice_ref.stable_dt(h, 4500.0, order), rel=1e-14)


def _csr(g, nodes):
    nodes = np.asarray(nodes, np.int64)
    nrows = _lib.load().wfd_points_nrows(ctypes.byref(g))
    rp = np.zeros(nrows + 1, np.int32)
    col = np.zeros(max(len(nodes), 1), np.int32)
    ids = np.zeros(max(len(nodes), 1), np.int32)
    rows = np.zeros(max(len(nodes), 1), np.int32)
    i32 = ctypes.POINTER(ctypes.c_int32)
    rc = _lib.load().wfd_points_csr(ctypes.byref(g), len(nodes), nodes.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                    rp.ctypes.data_as(i32), col.ctypes.data_as(i32), ids.ctypes.data_as(i32),
                                    rows.ctypes.data_as(i32))
    if rc == 0:  # the row array repeats the CSR row of each point
        for r in range(len(rp) - 1):
            assert all(rows[rp[r]:rp[r + 1]] == r)
    return rc, rp, col[: len(nodes)], ids[: len(nodes)]


def test_points_csr_sorted_and_stable():
    rc, g = _grid((20, 30), (1.0, 1.0))
    pts = [(5, 7), (3, 9), (5, 2), (5, 7), (18, 28), (3, 1)]
    flat = [i * 30 + j for i, j in pts]
    rc, rp, col, ids = _csr(g, flat)
    assert rc == 0 and rp[-1] == len(pts)
    for r in range(20):
        seg = slice(rp[r], rp[r + 1])
        assert all(flat[k] // 30 == r for k in ids[seg])
        assert list(col[seg]) == sorted(col[seg])
    # duplicates keep the caller's order
    dup = [k for k in ids if flat[k] == 5 * 30 + 7]
    assert dup == [0, 3]
    # row ranges and columns reproduce the points
    rebuilt = sorted((r, int(c), int(k)) for r in range(20) for c, k in zip(col[rp[r]:rp[r + 1]], ids[rp[r]:rp[r + 1]]))
    assert sorted((p[0], p[1], k) for k, p in enumerate(pts)) == rebuilt


def test_points_csr_3d_rows():
    rc, g = _grid((6, 7, 8), (1.0, 1.0, 1.0))
    pts = [(2, 3, 4), (1, 1, 1), (4, 5, 6)]
    flat = [np.ravel_multi_index(p, (6, 7, 8)) for p in pts]
    rc, rp, col, ids = _csr(g, flat)
    assert rc == 0 and len(rp) == 6 * 7 + 1
    assert rp[2 * 7 + 3 + 1] - rp[2 * 7 + 3] == 1 and col[rp[2 * 7 + 3]] == 4


@pytest.mark.parametrize("node", [(0, 5), (19, 5), (5, 0), (5, 29), (25, 3)])
def test_points_csr_rejects_non_interior(node):
    rc, g = _grid((20, 30), (1.0, 1.0))
    flat = node[0] * 30 + node[1] if node[0] < 20 else 20 * 30 + 3
    rc, *_ = _csr(g, [flat])
    assert rc == _lib.WFD_EINVAL


def test_step_argument_validation_needs_no_gpu():
    lib = _lib.load()
    assert lib.wfd_step(None, None, None, None, None, None, None, 0, None) == _lib.WFD_EINVAL
    rc, g = _grid((20, 30), (1.0, 1.0))
    med = _lib.Medium()
    med.kconst = 0.1
    buf = (ctypes.c_float * 4)()
    # u_next aliasing u_cur is rejected before any launch
    p = ctypes.addressof(buf)
    assert lib.wfd_step(ctypes.byref(g), p + 64, p, p, ctypes.byref(med), None, None, 0, None) == _lib.WFD_EINVAL
    med.kconst = float("nan")
    assert lib.wfd_step(ctypes.byref(g), p, p + 64, p, ctypes.byref(med), None, None, 0, None) == _lib.WFD_ENONFINITE
    assert lib.wfd_plan_create(ctypes.byref(ctypes.c_void_p()), ctypes.byref(g), 0, 0, 0) == _lib.WFD_EINVAL


@pytest.mark.parametrize("shape", [(17, 23), (9, 11, 13), (31,)])
def test_layout_views_round_trip_on_cpu(shape):
    g = solver.WaveGrid(shape, [1.0] * len(shape), 0.1, 4, torch.float64, nshots=2, device="cpu")
    a = torch.randn(2, *shape, dtype=torch.float64)
    f = g.from_nodes(a)
    assert f.shape == (2, g.elems)
    assert torch.equal(g.nodes(f), a)
    mask = torch.ones_like(f, dtype=torch.bool)
    g.nodes(mask).fill_(False)
    assert (f[mask] == 0).all()
    assert g.interior_cells == int(np.prod([n - 2 for n in shape]))


def _ac_medium():
    m = _lib.AcMedium()
    m.k = m.cn = m.cw = 0x1000  # never dereferenced: every call below fails validation first
    return m


@pytest.mark.parametrize("case", ["null_grid", "null_medium", "ndim3", "bad_order", "alias", "neg_steps",
                                  "null_coef", "bad_points"])
def test_acoustic_entry_points_validate_arguments(case):
    lib = _lib.load()
    shape = (6, 7, 8) if case == "ndim3" else (12, 14)
    rc, g = _grid(shape, [1.0] * len(shape), order=2)
    assert rc == 0
    if case == "bad_order":
        g.order = 5
    med = _ac_medium()
    if case == "null_coef":
        med.cn = None
    a, b = ctypes.c_void_p(0x2000), ctypes.c_void_p(0x3000 if case != "alias" else 0x2000)
    pts = None
    if case == "bad_points":
        p = _lib.Points()
        p.count, p.nrows = 1, 3  # the grid has 12 CSR rows
        p.row_ptr = p.col = p.id = p.row = p.data = 0x4000
        p.width, p.nsteps, p.shot_stride = 1, 4, 4
        pts = ctypes.byref(p)
    gp = None if case == "null_grid" else ctypes.byref(g)
    mp = None if case == "null_medium" else ctypes.byref(med)
    steps = -1 if case == "neg_steps" else 4
    rc = lib.wfd_ac_forward(gp, a, b, mp, pts, None, None, 0, steps, None)
    assert rc == _lib.WFD_EINVAL, _lib.last_error()
    c, d = ctypes.c_void_p(0x5000), ctypes.c_void_p(0x6000)
    rc = lib.wfd_ac_backward(gp, a, b, c, d, mp, pts, None, None, None, steps, ctypes.c_void_p(0x7000),
                             ctypes.c_void_p(0x8000), None)
    assert rc == _lib.WFD_EINVAL, _lib.last_error()
    assert _lib.last_error()


def test_launch_config_rules_need_no_gpu():
    """Launch shapes are host state: the streaming bit (WFD_VARIANT_STREAM = 2) is accepted on both
    2D halo variants and the 3D kernel, the shot-order bit (WFD_VARIANT_SHOTS = 4) on 2D grids only,
    anything else is rejected, and an untuned grid streams
    exactly when its four step fields exceed 512 MiB."""
    lib = _lib.load()
    v, r = ctypes.c_int32(), ctypes.c_int32()

    def cfg(g):
        assert lib.wfd_get_config(ctypes.byref(g), ctypes.byref(v), ctypes.byref(r)) == 0
        return v.value, r.value

    _, small = _grid((2048, 2048), (10.0, 10.0))
    _, big = _grid((8192, 8192), (10.0, 10.0))
    _, big64 = _grid((6000, 6000), (10.0, 10.0), dtype=_lib.WFD_F64)
    _, batch = _grid((2048, 2048), (10.0, 10.0), nshots=16)
    _, fwi = _grid((1024, 1024), (10.0, 10.0), nshots=16)
    _, cube = _grid((512, 512, 512), (10.0, 10.0, 10.0))
    _, small3 = _grid((128, 128, 128), (10.0, 10.0, 10.0))
    assert cfg(small)[0] & 2 == 0 and cfg(big)[0] & 2 == 2 and cfg(big64)[0] & 2 == 2
    assert cfg(batch)[0] & 2 == 2  # 16 shots of 2048^2: 16 x 4 fields x 17 MB
    assert cfg(fwi)[0] & 2 == 0  # 16 shots of 1024^2: 279 MB, in good part cache-resident
    # several shots take their tiles shot-fastest (WFD_VARIANT_SHOTS = 4) unless tuned otherwise
    assert cfg(fwi)[0] & 4 == 4 and cfg(batch)[0] & 4 == 4 and cfg(small)[0] & 4 == 0 and cfg(cube)[0] & 4 == 0
    assert cfg(cube)[0] == 2 and cfg(small3)[0] == 0
    for variant in range(8):  # bit 2: shot-fastest tile order (WFD_VARIANT_SHOTS)
        assert lib.wfd_set_config(ctypes.byref(small), variant, 4) == 0
        assert cfg(small) == (variant, 4)
    for variant in (-1, 8, 9):
        assert lib.wfd_set_config(ctypes.byref(small), variant, 4) == _lib.WFD_EINVAL
    assert lib.wfd_set_config(ctypes.byref(cube), 2, 16) == 0 and cfg(cube) == (2, 16)
    for variant in (1, 3, 4, 6):
        assert lib.wfd_set_config(ctypes.byref(cube), variant, 16) == _lib.WFD_EINVAL
    for g in (small, cube):
        assert lib.wfd_set_config(ctypes.byref(g), 0, 0) == 0  # forget: back to the default
    assert cfg(small)[0] & 2 == 0 and cfg(cube) == (2, 32)


def test_measurement_and_two_step_arguments_need_no_gpu():
    """wfd_stream_bench and wfd_forward2 reject bad arguments before any launch."""
    lib = _lib.load()
    buf = (ctypes.c_float * 64)()
    p = ctypes.addressof(buf)
    p16 = (p + 15) // 16 * 16
    assert lib.wfd_stream_bench(p16, 0, 0, p16, 8, 7, None) == _lib.WFD_EINVAL  # mode
    assert lib.wfd_stream_bench(p16, 0, 0, p16, 8, 1, None) == _lib.WFD_EINVAL  # b, c missing
    assert lib.wfd_stream_bench(p16, 0, 0, p16, 6, 0, None) == _lib.WFD_EINVAL  # n % 4
    assert lib.wfd_stream_bench(p16 + 4, 0, 0, p16, 8, 0, None) == _lib.WFD_EINVAL  # alignment
    assert lib.wfd_stream_bench(0, 0, 0, p16, 8, 0, None) == _lib.WFD_EINVAL
    assert lib.wfd_stream_bench(p16, 0, 0, p16, 0, 0, None) == _lib.WFD_OK  # nothing to do
    rc, g = _grid((20, 30), (1.0, 1.0))
    med = _lib.Medium()
    med.kconst = 0.1
    f = [p16 + 64 * i for i in range(4)]
    for rows in (1, 3, 5, 64):
        assert lib.wfd_forward2(ctypes.byref(g), f[0], f[1], f[2], f[3], ctypes.byref(med), None, None, 0, 4, rows,
                                None) == _lib.WFD_EINVAL
        assert "rows" in _lib.last_error()
    _, g3 = _grid((9, 11, 13), (1.0, 1.0, 1.0))
    for planes in (-1, 4097):  # 3D: rows = planes per block
        assert lib.wfd_forward2(ctypes.byref(g3), f[0], f[1], f[2], f[3], ctypes.byref(med), None, None, 0, 4, planes,
                                None) == _lib.WFD_EINVAL
        assert "planes" in _lib.last_error()
    _, g1 = _grid((40,), (1.0,))
    assert lib.wfd_forward2(ctypes.byref(g1), f[0], f[1], f[2], f[3], ctypes.byref(med), None, None, 0, 4, 0,
                            None) == _lib.WFD_EINVAL


def test_acoustic_rows_override_validation():
    lib = _lib.load()
    for bad in (-1, 3, 5, 16):
        assert lib.wfd_ac_set_rows(bad) == _lib.WFD_EINVAL
    for ok in (1, 2, 4, 8, 0):
        assert lib.wfd_ac_set_rows(ok) == _lib.WFD_OK


@pytest.mark.parametrize("ndim,order", [(1, 2), (2, 2), (2, 8), (3, 4), (3, 8)])
def test_cfl_limit_of_constant_coefficient_needs_no_gpu(ndim, order):
    """A constant kc above 4 / sum_d(S / h_d^2) (the bound wfd_stable_dt states) is rejected with
    WFD_EINVAL by every stream-ordered entry point before any launch; at the limit it passes the
    host check.  The limit is the closed form of the stencil symbol, checked here independently
    from the exact rational weights."""
    from fractions import Fraction
    from math import factorial

    lib = _lib.load()
    h = [0.7, 1.3, 2.1][:ndim]
    rc, g = _grid([12] * ndim, h, order)
    assert rc == 0
    K = order // 2
    a = [Fraction(2 * (-1) ** (k + 1) * factorial(K) ** 2, k * k * factorial(K - k) * factorial(K + k))
         for k in range(1, K + 1)]
    S = 2 * abs(sum(a)) + 2 * sum(abs(x) for x in a)  # |a_0| + 2 sum |a_k|, a_0 = -2 sum a_k
    limit = 4.0 / sum(float(S) / (x * x) for x in h)
    # the same limit as wfd_stable_dt: kc = (vmax dt)^2
    hh = (ctypes.c_double * 3)(*h, *([1.0] * (3 - ndim)))
    assert abs(lib.wfd_stable_dt(ndim, hh, 1.0, order) ** 2 / limit - 1) < 1e-12
    buf = (ctypes.c_float * 256)()
    p = (ctypes.addressof(buf) + 15) // 16 * 16
    med = _lib.Medium()
    flag = ctypes.c_int32()
    for kc, want in ((limit * 1.001, _lib.WFD_EINVAL), (limit * 1.0000001, None)):
        med.kconst = kc
        rc = lib.wfd_check_medium_async(ctypes.byref(g), ctypes.byref(med), ctypes.byref(flag), None)
        if want is None:
            continue  # inside the slack: the host check passes (the call goes on to the device)
        assert rc == want and "CFL" in _lib.last_error()
        assert lib.wfd_step(ctypes.byref(g), p, p + 256, p, ctypes.byref(med), None, None, 0, None) == want
        assert lib.wfd_forward(ctypes.byref(g), p, p + 256, ctypes.byref(med), None, None, 0, 4, None) == want
        assert lib.wfd_forward_graph(ctypes.byref(g), p, p + 256, ctypes.byref(med), None, None, 0, 40, 20,
                                     None) == want
        fg = ctypes.c_void_p()
        assert lib.wfd_fgraph_create(ctypes.byref(fg), ctypes.byref(g), p, p + 256, ctypes.byref(med), None, None,
                                     20) == want
        assert not fg.value
        assert lib.wfd_check_medium(ctypes.byref(g), ctypes.byref(med), None) == want


def test_graph_object_argument_validation_needs_no_gpu():
    lib = _lib.load()
    rc, g = _grid((20, 30), (1.0, 1.0))
    med = _lib.Medium()
    med.kconst = 0.1
    buf = (ctypes.c_float * 128)()
    p = (ctypes.addressof(buf) + 15) // 16 * 16
    fg = ctypes.c_void_p()
    for chunk in (0, 1, 3, 21):
        assert lib.wfd_fgraph_create(ctypes.byref(fg), ctypes.byref(g), p, p + 256, ctypes.byref(med), None, None,
                                     chunk) == _lib.WFD_EINVAL
        assert "chunk" in _lib.last_error() and not fg.value
    assert lib.wfd_fgraph_create(None, ctypes.byref(g), p, p + 256, ctypes.byref(med), None, None, 20) == _lib.WFD_EINVAL
    assert lib.wfd_fgraph_run(None, 0, 10, None) == _lib.WFD_EINVAL
    assert lib.wfd_fgraph_destroy(None) == _lib.WFD_OK


def test_noalloc_adjoint_rejects_problems_beyond_its_scratch_needs_no_gpu():
    """wfd_ac_backward_noalloc refuses, before any device work, a problem whose work fields
    ((2 + 2K+1) fields of nshots * elems, each rounded to 256 bytes) exceed WFD_AC_SCRATCH_BYTES."""
    lib = _lib.load()
    assert _lib.WFD_AC_SCRATCH_BYTES == 64 << 20
    buf = (ctypes.c_float * 256)()
    p = (ctypes.addressof(buf) + 255) // 256 * 256
    med = _lib.AcMedium()
    med.k = med.cn = med.cw = p
    rc, g = _grid((1024, 1024), (1.0, 1.0), order=8, nshots=8)
    need = 11 * ((g.elems * 8 * 4 + 255) // 256 * 256)
    assert need > _lib.WFD_AC_SCRATCH_BYTES
    G = (ctypes.c_double * 4)()
    rc = lib.wfd_ac_backward_noalloc(ctypes.byref(g), p, p + 64, ctypes.byref(med), None, None, None, None, 4,
                                     ctypes.addressof(G), None)
    assert rc == _lib.WFD_EINVAL and "at most" in _lib.last_error()
    assert lib.wfd_ac_backward_noalloc(ctypes.byref(g), p, p, ctypes.byref(med), None, None, None, None, 4,
                                       ctypes.addressof(G), None) == _lib.WFD_EINVAL  # aliased state
