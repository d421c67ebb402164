"""GPU forward parity: the HIP kernels against the C oracle (bit-exact, same operation order) and
against the slice-style restatement of the reference (tolerance), through the C ABI."""

import numpy as np
import pytest

from oracle import slice_ref
from tests.wave_cases import ricker_256_order2_setup, gpu_objects, make_case, oracle_forward, rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _run_gpu(case, tdt, nshots=1, graph_chunk=0):
    from neuralmech_amd import solver

    g, med, src, rec = gpu_objects(case, tdt, nshots=nshots)
    tr, (up, uc) = solver.run(g, med, case["nsteps"], src=src, sig=case["sig"] if src else None, rec=rec,
                              graph_chunk=graph_chunk)
    torch.cuda.synchronize()
    tr = None if tr is None else tr.cpu().numpy()
    return tr, g.nodes(up).cpu().numpy(), g.nodes(uc).cpu().numpy()


CASES = [
    # (shape, order, varc, sponge)
    ((37, 53), 2, False, False),
    ((37, 53), 2, True, True),
    ((29, 70), 4, True, False),
    ((41, 45), 6, False, True),
    ((33, 61), 8, True, True),
    ((300, 290), 8, True, True),
    ((53,), 4, True, True),
    ((101,), 8, False, False),
    ((13, 17, 70), 2, True, True),
    ((21, 19, 35), 8, True, True),
    ((24, 23, 22), 4, False, False),
    ((18, 40, 131), 6, True, True),
]


@pytest.mark.parametrize("dtype", ["f32", "f64"])
@pytest.mark.parametrize("shape,order,varc,sponge", CASES)
def test_forward_bitexact_vs_c_oracle(shape, order, varc, sponge, dtype):
    print(f"\n[forward] shape={shape} order={order} varc={varc} sponge={sponge} {dtype}", flush=True)
    npdt, tdt = (np.float32, torch.float32) if dtype == "f32" else (np.float64, torch.float64)
    case = make_case(shape, order, seed=hash((shape, order)) % 1000, varc=varc, sponge=sponge, nsteps=37)
    tr_o, up_o, uc_o = oracle_forward(case, npdt)
    tr_g, up_g, uc_g = _run_gpu(case, tdt)
    assert np.isfinite(uc_o).all() and np.abs(uc_o).max() > 0
    assert np.array_equal(uc_g[0], uc_o), f"u_cur differs: rel {rel_l2(uc_g[0], uc_o):.3e}"
    assert np.array_equal(up_g[0], up_o), f"u_prev differs: rel {rel_l2(up_g[0], up_o):.3e}"
    assert np.array_equal(tr_g[0], tr_o)


def test_ricker_256_order2_1000_steps_vs_slice_form_fp32():
    """256^2, order 2, 1000 steps, fp32: wavefield and 64 traces within 1e-5 relative L2 of the
    slice-form restatement of the reference step, and bit-exact with the C oracle."""
    case = ricker_256_order2_setup()
    print("\n[ricker 256x256] order 2 1000 steps fp32", flush=True)
    tr_g, up_g, uc_g = _run_gpu(case, torch.float32)
    kc32 = np.full(case["shape"], case["kconst"], np.float32)
    tr_s, _, uc_s = slice_ref.forward(kc32, 2, case["h"], case["nsteps"], None, case["src"],
                                      case["sig"].astype(np.float32), case["rec"])
    tr_64, _, uc_64 = slice_ref.forward(np.full(case["shape"], case["kconst"]), 2, case["h"], case["nsteps"],
                                        None, case["src"], case["sig"], case["rec"])
    e_field = rel_l2(uc_g[0], uc_s)
    e_tr = rel_l2(tr_g[0], tr_s)
    print(f"fp32 vs reference-slice fp32: field {e_field:.2e} traces {e_tr:.2e}; "
          f"vs fp64: field {rel_l2(uc_g[0], uc_64):.2e} traces {rel_l2(tr_g[0], tr_64):.2e}")
    assert e_field < 1e-5 and e_tr < 1e-5
    tr_o, _, uc_o = oracle_forward(case, np.float32)
    assert np.array_equal(uc_g[0], uc_o) and np.array_equal(tr_g[0], tr_o)


def test_fp64_gpu_matches_golden_fixture():
    """The 256^2 order-2 Ricker shot in float64 on the GPU against the committed fixture of the
    float64 slice-form restatement (tests/golden/ricker256_order2.npz)."""
    from pathlib import Path

    gold = np.load(Path(__file__).parent / "golden" / "ricker256_order2.npz")
    case = ricker_256_order2_setup()
    tr_g, _, uc_g = _run_gpu(case, torch.float64)
    assert rel_l2(tr_g[0][::10], gold["traces_every_10"]) < 1e-12
    assert rel_l2(uc_g[0].sum(1), gold["row_sums"]) < 1e-12 and rel_l2(uc_g[0].sum(0), gold["col_sums"]) < 1e-12
    assert abs(np.linalg.norm(uc_g[0]) / float(gold["field_l2"]) - 1) < 1e-12


def test_fp64_gpu_matches_reference_slice_form():
    case = make_case((64, 80), 8, seed=5, nsteps=200)
    tr_g, _, uc_g = _run_gpu(case, torch.float64)
    beta = slice_ref.beta_field(case["prof"], np.float64)
    tr_s, _, uc_s = slice_ref.forward(case["kc"], 8, case["h"], 200, beta, case["src"], case["sig"], case["rec"])
    assert rel_l2(uc_g[0], uc_s) < 1e-12 and rel_l2(tr_g[0], tr_s) < 1e-12


def test_batched_shots_equal_single_shots():
    from neuralmech_amd import solver

    case = make_case((45, 77), 8, seed=11, nsteps=30)
    g, med, _, _ = gpu_objects(case, torch.float32, nshots=3)
    srcs = [[(10, 10)], [(20, 30), (22, 31)], [(30, 60)]]
    recs = [[(5, j) for j in range(1, 76, 5)]] * 3
    sig = np.random.default_rng(0).standard_normal((3, 30, 2))
    src = solver.PointSet(g, srcs, per_shot=True)
    rec = solver.PointSet(g, recs, per_shot=True)
    tr, (_, uc) = solver.run(g, med, 30, src=src, sig=torch.as_tensor(sig), rec=rec)
    for s in range(3):
        c1 = dict(case, src=srcs[s], rec=recs[s], sig=sig[s][:, : len(srcs[s])])
        tr_o, _, uc_o = oracle_forward(c1, np.float32)
        assert np.array_equal(g.nodes(uc)[s].cpu().numpy(), uc_o)
        assert np.array_equal(tr[s].cpu().numpy(), tr_o)


def test_graph_replay_bitexact_with_plain_launches():
    case = make_case((70, 90), 4, seed=3, nsteps=47)
    a = _run_gpu(case, torch.float32)
    b = _run_gpu(case, torch.float32, graph_chunk=10)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_in_place_and_separate_output_agree():
    from neuralmech_amd import _lib, solver

    case = make_case((40, 50), 6, seed=8, nsteps=1, nsrc=0, nrec=0)
    g, med, _, _ = gpu_objects(case, torch.float32)
    rng = np.random.default_rng(1)
    a, b = rng.standard_normal((2, *case["shape"]))
    for x in (a, b):  # boundary nodes hold zero by the layout contract
        x[0, :] = x[-1, :] = x[:, 0] = x[:, -1] = 0.0
    up = g.from_nodes(torch.as_tensor(a).float())
    uc = g.from_nodes(torch.as_tensor(b).float())
    out = g.zeros()
    lib = _lib.load()
    import ctypes

    st = solver._stream_ptr(g.device)
    _lib.check(lib.wfd_step(ctypes.byref(g.c), up.data_ptr(), uc.data_ptr(), out.data_ptr(), ctypes.byref(med.c),
                            None, None, 0, st))
    up2 = up.clone()
    _lib.check(lib.wfd_step(ctypes.byref(g.c), up2.data_ptr(), uc.data_ptr(), up2.data_ptr(), ctypes.byref(med.c),
                            None, None, 0, st))
    torch.cuda.synchronize()
    assert torch.equal(out, up2)
    # halo and padding stay zero
    mask = torch.ones_like(out, dtype=torch.bool)
    g.nodes(mask).fill_(False)
    assert (out[mask] == 0).all()


def test_aliasing_and_bad_points_rejected():
    from neuralmech_amd import _lib, solver

    case = make_case((20, 20), 2, seed=1, nsteps=1, nsrc=0, nrec=0)
    g, med, _, _ = gpu_objects(case, torch.float32)
    u = g.zeros()
    import ctypes

    rc = _lib.load().wfd_step(ctypes.byref(g.c), u.data_ptr(), u.data_ptr(), u.data_ptr(), ctypes.byref(med.c),
                              None, None, 0, solver._stream_ptr(g.device))
    assert rc == _lib.WFD_EINVAL
    with pytest.raises(_lib.WaveError):
        solver.PointSet(g, [(0, 5)])  # boundary node


def test_medium_check_reports_nonfinite():
    from neuralmech_amd import _lib, solver

    g = solver.WaveGrid((30, 30), (1.0, 1.0), 0.1, 4)
    v = torch.ones(30, 30, dtype=torch.float64)
    solver.Medium(g, velocity=v).validate()
    v[7, 9] = float("nan")
    with pytest.raises(_lib.WaveError) as e:
        solver.Medium(g, velocity=v).validate()
    assert e.value.code == _lib.WFD_ENONFINITE


# (variant, rows): variant bit 0 = lane-exchange halo, bit 1 = streaming access to u_prev, kc, u_next
SHAPES_2D = [(v, r) for v in (0, 1, 2, 3) for r in (1, 2, 4, 8)]


@pytest.mark.parametrize("variant,rows", SHAPES_2D)
def test_every_2d_launch_shape_bitexact(variant, rows):
    from neuralmech_amd import solver

    print(f"\n[launch] 2D variant={variant} rows={rows}", flush=True)
    for shape, order in (((61, 530), 8), ((70, 249), 4), ((33, 300), 2)):
        case = make_case(shape, order, seed=7 + order, nsteps=25)
        g, med, src, rec = gpu_objects(case, torch.float32)
        g.set_launch(variant, rows)
        assert g.launch() == (variant, rows)
        try:
            tr, (_, uc) = solver.run(g, med, 25, src=src, sig=case["sig"], rec=rec)
            torch.cuda.synchronize()
        finally:
            g.set_launch(0, 0)
        tr_o, _, uc_o = oracle_forward(case, np.float32)
        assert np.array_equal(g.nodes(uc)[0].cpu().numpy(), uc_o), (shape, order)
        assert np.array_equal(tr[0].cpu().numpy(), tr_o)


@pytest.mark.parametrize("variant,rows", [(0, 1), (0, 3), (0, 4), (0, 8), (0, 16), (0, 64), (2, 4), (2, 7), (2, 32)])
def test_every_3d_launch_shape_bitexact(variant, rows):
    from neuralmech_amd import solver

    print(f"\n[launch] 3D variant={variant} rows={rows}", flush=True)
    # the second grid spans several tile rows and columns (tile boundaries inside the grid)
    for shape, seed in (((30, 21, 70), 31), ((12, 75, 140), 37)):
        case = make_case(shape, 8, seed=seed, nsteps=12)
        g, med, src, rec = gpu_objects(case, torch.float32)
        g.set_launch(variant, rows)
        try:
            tr, (_, uc) = solver.run(g, med, 12, src=src, sig=case["sig"], rec=rec)
            torch.cuda.synchronize()
        finally:
            g.set_launch(0, 0)
        tr_o, _, uc_o = oracle_forward(case, np.float32)
        assert np.array_equal(g.nodes(uc)[0].cpu().numpy(), uc_o), shape
        assert np.array_equal(tr[0].cpu().numpy(), tr_o), shape


@pytest.mark.parametrize("shape,variant,rows", [((61, 530), 3, 4), ((61, 530), 2, 4), ((30, 21, 70), 2, 8)])
def test_streaming_access_fp64_bitexact(shape, variant, rows):
    """The streaming variants in float64 (two 16-byte hinted accesses per 4-node chunk), in-place
    leapfrog (u_next aliases u_prev), against the C oracle bit for bit."""
    from neuralmech_amd import solver

    print(f"\n[launch] fp64 streaming {shape} variant={variant} rows={rows}", flush=True)
    case = make_case(shape, 8, seed=43, nsteps=15)
    g, med, src, rec = gpu_objects(case, torch.float64)
    g.set_launch(variant, rows)
    try:
        tr, (_, uc) = solver.run(g, med, 15, src=src, sig=case["sig"], rec=rec)
        torch.cuda.synchronize()
    finally:
        g.set_launch(0, 0)
    tr_o, _, uc_o = oracle_forward(case, np.float64)
    assert np.array_equal(g.nodes(uc)[0].cpu().numpy(), uc_o)
    assert np.array_equal(tr[0].cpu().numpy(), tr_o)


def test_tuner_picks_a_valid_shape_and_results_do_not_change():
    from neuralmech_amd import solver

    case = make_case((257, 700), 8, seed=41, nsteps=20)
    g, med, src, rec = gpu_objects(case, torch.float32)
    info = g.tune(med, steps=6)
    assert (info["variant"], info["rows"]) in SHAPES_2D and info["us_per_step"] > 0
    tr, (_, uc) = solver.run(g, med, 20, src=src, sig=case["sig"], rec=rec)
    torch.cuda.synchronize()
    g.set_launch(0, 0)
    tr_o, _, uc_o = oracle_forward(case, np.float32)
    assert np.array_equal(g.nodes(uc)[0].cpu().numpy(), uc_o)
