"""The allocation-free adjoint wfd_ac_backward_noalloc (work fields in the library's per-device
scratch array) against wfd_ac_backward (work fields passed by the caller), through the C ABI:
bit-identical sums G and source records on two calls in a row, the second on a smaller problem
(so the scratch holds the first call's values when it starts), and a clear error for a problem
whose work fields exceed the array (tests/test_abi_host.py)."""

import numpy as np
import pytest

from tests.test_gpu_acoustic import _case

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _both(C, dJ):
    """Forward once, then the adjoint with caller work fields and with the scratch array, each on
    its own copy of the final state (the adjoint consumes it)."""
    ac = C["ac"]
    sig = torch.as_tensor(C["sig"])
    tr, store, state = ac.forward(C["grid"], C["med"], C["N"], C["src"], sig, C["rec"], C["strip"])
    out = {}
    for noalloc in (True, False):
        st = state.clone()
        G, mu = ac.backward(C["grid"], C["med"], C["N"], st, C["src"], sig, C["rec"], dJ(tr), C["strip"], store,
                            noalloc=noalloc)
        torch.cuda.synchronize()
        out[noalloc] = (G, mu)
    return out


def test_noalloc_adjoint_equals_workspace_adjoint_on_two_calls():
    from neuralmech_amd import _lib

    assert _lib.WFD_AC_SCRATCH_BYTES == 64 << 20
    print("\n[noalloc] call 1: 2 shots 150 x 170 order 4, 90 steps", flush=True)
    big = _case(torch.float32, "neumann", shape=(150, 170), N=90, order=4, seed=7)
    r1 = _both(big, lambda tr: 2.0 * tr.double())
    print("[noalloc] call 2: 2 shots 37 x 70 order 2, 64 steps", flush=True)
    small = _case(torch.float32, "dirichlet", shape=(37, 70), N=64, order=2, seed=3)
    r2 = _both(small, lambda tr: tr.double() ** 3)
    for r in (r1, r2):
        (Gn, mun), (Gw, muw) = r[True], r[False]
        assert torch.isfinite(Gw).all() and Gw.abs().max() > 0
        assert torch.equal(Gn, Gw), "sums differ between the scratch array and caller work fields"
        assert torch.equal(mun, muw), "source records differ"
