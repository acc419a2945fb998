"""The request loop (hydraphase.serve): continuous batching under an HBM
budget.  Requests arrive faster than they finish, so later ones join while
earlier ones run; the budget is set so that not all fit at once.  Every
result must equal the same case run alone, bit for bit (FP64 and FP32-bulk
alike: a case's arithmetic never reads another case)."""
import numpy as np
import pytest

from hydraphase import serve as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,policy", [(0, "fifo"), (1, "fifo"), (0, "shortest")])
def test_served_results_equal_alone(precision, policy):
    """Smallest-first admission with a fill target changes only when each
    case runs and beside which others."""
    reqs = S.poisson_requests(8, rate=400.0, seed=3, nmin=4, nmax=8, steps=(3, 12), precision=precision)
    biggest = max(r.nbytes for r in reqs)
    budget = 2.5 * biggest
    fill = None if policy == "fifo" else 2 * max(int(r.mesh.iblank.sum()) for r in reqs)
    srv = S.Server(budget_bytes=budget, max_cases=16, chunk=4, policy=policy, fill_cells=fill)
    try:
        stats = srv.run(reqs)
    finally:
        srv.close()
    print(stats)
    assert stats["peak_resident_bytes"] <= budget
    # some requests had to wait for memory, and some ran side by side
    assert any(r.admit > r.arrival + 1e-4 for r in reqs)
    overlap = any(a.admit < b.finish and b.admit < a.finish for a in reqs for b in reqs if a is not b)
    assert overlap
    for r in reqs:
        ref = S.alone(r)
        cu, gp, t, tot = r.result
        assert tot == r.spec.steps == ref[3]
        assert t == ref[2]
        assert np.array_equal(cu, ref[0]) and np.array_equal(gp, ref[1]), r.rid
