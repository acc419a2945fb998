"""Host side of batching: the ensemble generator and the admission size
estimate (no GPU needed)."""
import math

import numpy as np

from hydraphase import batch as HB
from hydraphase import mesh as M


def test_ensemble_specs_distribution():
    sp = HB.ensemble_specs(64, seed=1)
    n = np.array([s.n for s in sp])
    assert len(sp) == 64 and n.min() >= 8 and n.max() <= 48
    # log-uniform: about half the cases below the geometric mean of the range
    gm = math.sqrt(8 * 49)
    assert 20 <= (n < gm).sum() <= 44
    assert sum(s.viscous for s in sp) == 32
    assert all(0.1 <= s.mach <= 0.8 and 0.05 <= s.radius <= 0.25 for s in sp)
    assert len({s.seed for s in sp}) == 64
    assert [s.n for s in HB.ensemble_specs(64, seed=1)] == list(n)


def test_case_bytes_scales_with_cells():
    small = M.box_mesh(4, 4, 4)
    big = M.box_mesh(8, 8, 8)
    b0, b1 = HB.case_bytes(small, 0), HB.case_bytes(big, 0)
    assert 0 < b0 < b1
    per_cell = (b1 - b0) / (big.neles - small.neles)
    assert 1000 < per_cell < 1700, per_cell  # about 1.5 KB of FP64 state and geometry per tet
    assert HB.case_bytes(big, 1) > b1      # FP32 gradients come on top
