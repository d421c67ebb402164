"""Write ``tests/golden/ricker256_order2.npz``: the 256 x 256, order-2 Ricker shot (dx = 1, c = 1,
c dt / dx = 0.5, 1000 steps, source at the centre, 64 receivers on row 20) in float64 by the
slice-form restatement of the reference update (``oracle/slice_ref.py``, following
``projects/18_simple_wave_fd/2D_wave_cupy.py:54-62``): every 10th trace row, and the final field's
row sums, column sums and L2 norm.

The reference ships no vectors for this path; this fixture pins the restatement against itself
over time (a changed oracle fails ``tests/test_oracle_pins.py``) and is the float64 target of the
GPU path.  Run from the repository root: ``python tests/golden/make_fixtures.py``.
"""

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import slice_ref  # noqa: E402
from tests.wave_cases import ricker_256_order2_setup  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ricker256_order2.npz")


def compute():
    case = ricker_256_order2_setup()
    tr, _, uc = slice_ref.forward(np.full(case["shape"], case["kconst"]), 2, case["h"], case["nsteps"], None,
                                  case["src"], case["sig"], case["rec"])
    return dict(traces_every_10=tr[::10].copy(), row_sums=uc.sum(1), col_sums=uc.sum(0),
                field_l2=np.array(np.linalg.norm(uc)))


if __name__ == "__main__":
    np.savez(OUT, **compute())
    print("wrote", OUT)
