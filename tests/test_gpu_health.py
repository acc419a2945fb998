"""Health checks (hp_health, hp_batch_health) and the request loop retiring
a failed case: a non-finite state is flagged for its own case only, and the
other requests of a mix still equal their alone runs bit for bit.  NaN is
ordinary data to every kernel (no kernel indexes memory by a state value)."""
import numpy as np
import pytest

from hydraphase import Solver
from hydraphase import _lib
from hydraphase import batch as HB
from hydraphase import mesh as M
from hydraphase import serve as S

pytestmark = pytest.mark.gpu


def small(seed=5):
    m = M.box_mesh(6, 5, 4, jitter=0.1, seed=seed)
    p = M.water_air_params(cfl=0.5, viscous_gas=True)
    cu, gp = M.sphere_state(m, p, [(0.0, 0.1, 0.0, 0.22)])
    return m, p, cu, gp


def test_healthy_case_reports_zero():
    m, p, cu, gp = small()
    s = Solver(m, p)
    try:
        s.upload(cu, gp, 0.0, 0)
        s.step(3)
        assert s.health() == 0
    finally:
        s.close()


def test_nonfinite_state_flagged():
    m, p, cu, gp = small()
    cu = cu.copy()
    cu[4, 17] = np.nan  # gas energy of one cell
    s = Solver(m, p)
    try:
        s.upload(cu, gp, 0.0, 0)
        s.step(1)
        assert s.health() & _lib.HEALTH_NONFINITE
    finally:
        s.close()


def test_pressure_flag_from_primitives():
    m, p, cu, gp = small()
    s = Solver(m, p)
    try:
        s.upload(cu, gp, 0.0, 0)
        s.step(1)
        q = s.get("q")
        q[3, 9] = -1.0  # pg of cell 10
        s.set("q", q)
        assert s.health() == _lib.HEALTH_PRESSURE
    finally:
        s.close()


def test_batch_flags_only_the_failed_slot():
    m, p, cu, gp = small()
    bad = cu.copy()
    bad[0, 3] = np.inf
    b = HB.Batch(4, 4 * m.neles, 4 * m.nghosts)
    sols = []
    try:
        slots = []
        for state in (cu, bad, cu):
            s = Solver(m, p)
            s.upload(state, gp, 0.0, 0)
            sols.append(s)
            slots.append(b.add(s))
        b.step(2, graph=False)
        h = b.health()
        assert h[slots[1]] & _lib.HEALTH_NONFINITE
        assert h[slots[0]] == 0 and h[slots[2]] == 0
    finally:
        b.close()
        for s in sols:
            s.close()


def test_served_mix_retires_failed_request():
    reqs = S.poisson_requests(5, rate=400.0, seed=7, nmin=4, nmax=7, steps=(6, 12), precision=0)
    victim = reqs[2]
    victim.cu = victim.cu.copy()
    victim.cu[4, 5] = np.nan
    srv = S.Server(budget_bytes=10 * max(r.nbytes for r in reqs), max_cases=8, chunk=3)
    try:
        stats = srv.run(reqs)
    finally:
        srv.close()
    print(stats)
    assert stats["failed"] == 1
    assert victim.error is not None and victim.result is None
    for r in reqs:
        if r is victim:
            continue
        assert r.error is None
        ref = S.alone(r)
        cu, gp, t, tot = r.result
        assert tot == r.spec.steps == ref[3] and t == ref[2]
        assert np.array_equal(cu, ref[0]) and np.array_equal(gp, ref[1]), r.rid


def test_reference_breakdown_flagged():
    """A water sphere (r = 0.2) in air on a 48^3 x 6 box: by step 10 the
    reference's own step drives the gas internal energy of three nearly
    gas-free cells negative, the mixture-pressure root takes the square root
    of a negative number and those cells become NaN (the oracle gives the
    same NaNs).  The health check flags that case, alone and in a batch,
    and not the healthy case beside it."""
    m = M.box_mesh(48, 48, 48, jitter=0.1, seed=1)
    p = M.water_air_params(cfl=0.5, viscous_gas=True)
    cu, gp = M.sphere_state(m, p, [(0.0, 0.1, 0.0, 0.2)])
    s = Solver(m, p)
    try:
        s.upload(cu, gp, 0.0, 0)
        s.step(9)
        assert s.health() == 0
        s.step(1)
        assert s.health() & _lib.HEALTH_NONFINITE
        assert np.isnan(s.download()[0]).any()
    finally:
        s.close()
    mh, ph, cuh, gph = small()
    b = HB.Batch(2, m.neles + mh.neles, m.nghosts + mh.nghosts)
    sols = []
    try:
        for mm, pp, cc, gg in ((m, p, cu, gp), (mh, ph, cuh, gph)):
            x = Solver(mm, pp)
            x.upload(cc, gg, 0.0, 0)
            sols.append((b.add(x), x))
        b.step(10, graph=True)
        h = b.health()
        assert h[sols[0][0]] & _lib.HEALTH_NONFINITE and h[sols[1][0]] == 0
    finally:
        b.close()
        for _, x in sols:
            x.close()
