"""CPU tests of the run-state contract: solver.WaveState is advanced in place (levels and step), a
plain (u_prev, u_cur) pair is refused, and the C oracle's chained runs equal one run (the
property the GPU resume tests check against the kernels)."""

import numpy as np
import pytest
import torch

from neuralmech_amd import solver
from oracle import c_oracle
from tests.wave_cases import make_case


def test_wave_state_reads_as_a_pair_and_checks_its_fields():
    a, b = torch.zeros(1, 8), torch.ones(1, 8)
    st = solver.WaveState(a, b, step=5)
    prev, cur = st
    assert prev is a and cur is b and st[1] is b and len(st) == 2 and st.step == 5
    c = st.clone()
    assert c.step == 5 and torch.equal(c.cur, b) and c.cur is not b
    with pytest.raises(ValueError):
        solver.WaveState(a, a)


def test_plain_pair_state_is_refused_and_step0_must_match():
    g = solver.WaveGrid((9, 11), (1.0, 1.0), 0.1, 2, torch.float32, device="cpu")
    with pytest.raises(TypeError, match="WaveState"):
        solver._as_state(g, (g.zeros(), g.zeros()), None)
    st = solver.WaveState.rest(g, step=7)
    assert solver._as_state(g, st, None) is st and solver._as_state(g, st, 7) is st
    with pytest.raises(ValueError, match="step 7"):
        solver._as_state(g, st, 3)
    fresh = solver._as_state(g, None, 4)
    assert fresh.step == 4 and not fresh.cur.any()


@pytest.mark.parametrize("n1", [5, 6, 1])
def test_oracle_chained_runs_equal_one_run(n1):
    """Resuming the C oracle from (u^{n1-1}, u^{n1}) with the signal rows from n1 on reproduces
    the uninterrupted run bit for bit, traces and fields, for odd and even n1."""
    case = make_case((31, 44), 8, seed=3, nsteps=25)
    kc = case["kc"].astype(np.float32)
    kw = dict(beta_profiles=case["prof"], src_nodes=case["src"], rec_nodes=case["rec"])
    sig = case["sig"].astype(np.float32)
    tr, up, uc = c_oracle.forward(kc, 8, case["h"], 25, src_sig=sig, **kw)
    t1, p1, c1 = c_oracle.forward(kc, 8, case["h"], n1, src_sig=sig[:n1], **kw)
    t2, p2, c2 = c_oracle.forward(kc, 8, case["h"], 25 - n1, src_sig=sig[n1:], u_prev=p1, u_cur=c1, **kw)
    assert np.array_equal(np.concatenate([t1, t2]), tr)
    assert np.array_equal(c2, uc) and np.array_equal(p2, up)
    # the mistake the resume contract rules out: continuing from the swapped levels
    _, _, bad = c_oracle.forward(kc, 8, case["h"], 25 - n1, src_sig=sig[n1:], u_prev=c1, u_cur=p1, **kw)
    assert not np.array_equal(bad, uc)
