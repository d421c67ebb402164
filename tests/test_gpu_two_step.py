"""GPU tests of the two-steps-per-launch 2D and 3D kernels: bit-identical to single steps (fields
and traces) over orders, sponge, sources, receivers, shapes narrower and wider than one tile,
planes per block (3D), odd and even step counts, batches of shots and both dtypes."""

import numpy as np
import pytest

from tests.wave_cases import gpu_objects, make_case, oracle_forward

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _run(case, dtype, nshots=1, two_step=0, nsteps=None, step0=0):
    from neuralmech_amd import solver

    g, med, src, rec = gpu_objects(case, dtype, nshots)
    N = case["nsteps"] if nsteps is None else nsteps
    sig = torch.as_tensor(case["sig"]) if src is not None else None
    tr, (up, uc) = solver.run(g, med, N, src=src, sig=sig, rec=rec, two_step=two_step)
    return tr, g.nodes(up), g.nodes(uc)


CASES = [((40, 50), 2, True, 13), ((70, 300), 8, True, 40), ((257, 33), 4, False, 21), ((33, 700), 6, True, 30),
         ((96, 241), 8, False, 17), ((120, 480), 8, True, 1), ((60, 60), 8, True, 2)]


@pytest.mark.parametrize("shape,order,sponge,nsteps", CASES)
@pytest.mark.parametrize("rows", [2, 4, 8, 16, 32])
def test_two_step_matches_single_steps(shape, order, sponge, nsteps, rows):
    case = make_case(shape, order, seed=sum(shape) + order, sponge=sponge, nsrc=3, nrec=9, nsteps=nsteps)
    print("case", shape, order, sponge, nsteps, rows, flush=True)
    tr1, up1, uc1 = _run(case, torch.float32)
    tr2, up2, uc2 = _run(case, torch.float32, two_step=rows)
    assert torch.equal(uc1, uc2) and torch.equal(up1, up2)
    assert torch.equal(tr1, tr2)


def test_two_step_matches_oracle_fp32_and_fp64():
    for dtype, npd in [(torch.float32, np.float32), (torch.float64, np.float64)]:
        case = make_case((90, 260), 8, seed=3, nsrc=2, nrec=7, nsteps=24)
        print("dtype", dtype, flush=True)
        for rows in (16, 8, 4):
            tr, up, uc = _run(case, dtype, two_step=rows)
            tr_o, up_o, uc_o = oracle_forward(case, npd)
            assert np.array_equal(uc[0].cpu().numpy(), uc_o) and np.array_equal(up[0].cpu().numpy(), up_o)
            assert np.array_equal(tr[0].cpu().numpy(), tr_o)


def test_two_step_batched_shots_and_constant_speed():
    case = make_case((64, 250), 4, seed=7, nsrc=2, nrec=5, nsteps=15)
    tr1, _, uc1 = _run(case, torch.float32, nshots=3)
    for rows in (8, 2):
        tr2, _, uc2 = _run(case, torch.float32, nshots=3, two_step=rows)
        assert torch.equal(uc1, uc2) and torch.equal(tr1, tr2)
    case = make_case((50, 90), 8, seed=8, varc=False, nsteps=12)
    tr1, _, uc1 = _run(case, torch.float32)
    tr2, _, uc2 = _run(case, torch.float32, two_step=16)
    assert torch.equal(uc1, uc2) and torch.equal(tr1, tr2)


CASES_3D = [((18, 40, 131), 8, True, 13, 3), ((30, 57, 70), 4, True, 10, 2), ((9, 9, 9), 2, False, 7, 1),
            ((40, 33, 130), 6, False, 9, 2), ((26, 61, 57), 8, True, 2, 3), ((21, 30, 200), 8, True, 1, 2)]


@pytest.mark.parametrize("shape,order,sponge,nsteps,nsrc", CASES_3D)
@pytest.mark.parametrize("planes", [1, 5, 64])
def test_two_step_3d_matches_single_steps(shape, order, sponge, nsteps, nsrc, planes):
    """Sources sit at random interior nodes, so with few planes per block they also fall in the
    planes that step 1 evaluates ahead of and behind a block's own planes."""
    case = make_case(shape, order, seed=sum(shape) + order, sponge=sponge, nsrc=nsrc, nrec=9, nsteps=nsteps)
    print("case", shape, order, sponge, nsteps, planes, flush=True)
    tr1, up1, uc1 = _run(case, torch.float32)
    tr2, up2, uc2 = _run(case, torch.float32, two_step=planes)
    assert torch.equal(uc1, uc2) and torch.equal(up1, up2)
    assert torch.equal(tr1, tr2)


def test_two_step_3d_matches_oracle_batches_and_constant_speed():
    from neuralmech_amd import _lib

    case = make_case((20, 30, 60), 8, seed=4, nsrc=2, nrec=7, nsteps=12)
    tr, up, uc = _run(case, torch.float32, two_step=8)
    tr_o, up_o, uc_o = oracle_forward(case, np.float32)
    assert np.array_equal(uc[0].cpu().numpy(), uc_o) and np.array_equal(up[0].cpu().numpy(), up_o)
    assert np.array_equal(tr[0].cpu().numpy(), tr_o)
    with pytest.raises(_lib.WaveError) as e:  # float32 only in 3D
        _run(case, torch.float64, two_step=8)
    assert e.value.code == _lib.WFD_EINVAL and "float32" in str(e.value)
    case = make_case((14, 50, 75), 4, seed=9, nsrc=2, nrec=5, nsteps=8)
    tr1, _, uc1 = _run(case, torch.float32, nshots=3)
    tr2, _, uc2 = _run(case, torch.float32, nshots=3, two_step=16)
    assert torch.equal(uc1, uc2) and torch.equal(tr1, tr2)
    case = make_case((12, 40, 70), 8, seed=10, varc=False, nsteps=6)
    tr1, _, uc1 = _run(case, torch.float32)
    tr2, _, uc2 = _run(case, torch.float32, two_step=4)
    assert torch.equal(uc1, uc2) and torch.equal(tr1, tr2)


def test_two_step_rejects_aliasing_and_1d():
    from neuralmech_amd import _lib, solver

    case = make_case((40, 40), 4, nsteps=4)
    g, med, _, _ = gpu_objects(case, torch.float32)
    a, b = g.zeros(), g.zeros()
    import ctypes

    lib = _lib.load()
    rc = lib.wfd_forward2(ctypes.byref(g.c), a.data_ptr(), b.data_ptr(), a.data_ptr(), b.data_ptr(),
                          ctypes.byref(med.c), None, None, 0, 4, 16, None)
    assert rc == _lib.WFD_EINVAL and "distinct" in _lib.last_error()
    g1 = solver.WaveGrid((50,), (1.0,), 0.1, 4, torch.float32)
    m1 = solver.Medium(g1, kconst=0.5)
    c, d, e, f = g1.zeros(), g1.zeros(), g1.zeros(), g1.zeros()
    rc = lib.wfd_forward2(ctypes.byref(g1.c), c.data_ptr(), d.data_ptr(), e.data_ptr(), f.data_ptr(),
                          ctypes.byref(m1.c), None, None, 0, 4, 0, None)
    assert rc == _lib.WFD_EINVAL and "3D" in _lib.last_error()
