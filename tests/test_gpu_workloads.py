"""The batched ensemble and the serving mix at their workload sizes.

* The uneven ensemble: 64 boxes of n^3 x 6 tets, n log-uniform 8..48
  (8,351,772 tets in all; seed 1), advanced 20 steps in one batch with the
  captured step graph, FP64: every case equals its alone hp_step run bit for
  bit, and the two largest (n = 47 and 46) equal the oracle's 20-step state
  by its hash (tests/golden/workload_hashes.json, written by
  tests/golden/make_workload_hashes.py).
* The serving mix: Poisson arrivals of boxes n 10..70 including an n = 70
  box (2,058,000 tets), admitted under an HBM budget that cannot hold them
  all, stepped together in chunks: every result equals its alone run bit for
  bit, and the n = 70 box equals the oracle's state by its hash.  A request
  whose state breaks down (the reference's own step turns some large water
  spheres non-finite within a few steps) is retired with an error by the
  loop's health check; its alone run must break down as well.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from hydraphase import Solver
from hydraphase import batch as HB
from hydraphase import serve as S

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
STEPS = 20


def hashes():
    return json.load(open(os.path.join(HERE, "golden", "workload_hashes.json")))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def alone(case, nsteps):
    m, p, cu, gp, _ = case
    s = Solver(m, p)
    s.upload(cu, gp, 0.0, 0)
    s.step(nsteps)
    out = s.download()
    s.close()
    return out


def same(a, b):
    return (a[3] == b[3] and a[2] == b[2] and np.array_equal(a[0], b[0], equal_nan=True)
            and np.array_equal(a[1], b[1], equal_nan=True))


def test_ensemble_at_size_batched_equals_alone():
    specs = HB.ensemble_specs(64, seed=1)
    cases = [HB.build_case(s, 0) for s in specs]
    print(f"ensemble: {sum(c[0].neles for c in cases)} tets in {len(cases)} cases", flush=True)
    b = HB.Batch(*HB.capacity([c[0] for c in cases]))
    sols = []
    for (m, p, cu, gp, _) in cases:
        s = Solver(m, p)
        s.upload(cu, gp, 0.0, 0)
        b.add(s)
        sols.append(s)
    b.step(STEPS, graph=True)
    b.synchronize()
    got = [s.download() for s in sols]
    b.close()
    for s in sols:
        s.close()
    for k, c in enumerate(cases):
        assert same(got[k], alone(c, STEPS)), (k, specs[k])
    for k, h in hashes()["ensemble"].items():
        k = int(k)
        assert specs[k].n == h["spec"][0] and got[k][3] == h["total"] == STEPS
        assert got[k][2] == h["time"]
        assert sha(got[k][0]) == h["cu_sha256"] and sha(got[k][1]) == h["gp_sha256"], (k, specs[k])


def test_serving_mix_at_size_equals_alone():
    reqs = S.poisson_requests(11, rate=20.0, seed=5, nmin=10, nmax=70, steps=(10, 30), precision=0)
    h = hashes()["serve_n70"]
    spec = HB.CaseSpec(*h["spec"], STEPS)
    m, p, cu, gp, _ = HB.build_case(spec, 0)
    big = S.Request(len(reqs), spec, 0.05, m, p, cu, gp, HB.case_bytes(m, 0))
    reqs.append(big)
    total = sum(r.nbytes for r in reqs)
    print(f"serving mix: {sum(r.mesh.neles for r in reqs)} tets in {len(reqs)} requests, "
          f"largest {big.mesh.neles}", flush=True)
    budget = 0.6 * total
    srv = S.Server(budget_bytes=budget, max_cases=16, chunk=5, policy="shortest")
    try:
        stats = srv.run(reqs)
    finally:
        srv.close()
    print(stats)
    assert stats["peak_resident_bytes"] <= budget
    assert any(r.admit > r.arrival + 1e-3 for r in reqs)       # some waited for memory
    assert any(a.admit < b.finish and b.admit < a.finish for a in reqs for b in reqs if a is not b)
    healthy = 0
    for r in reqs:
        ref = S.alone(r)
        if r.error is not None:
            assert not np.isfinite(ref[0]).all(), (r.rid, r.error)
            continue
        healthy += 1
        assert same(r.result, ref), r.rid
    assert healthy >= 8 and big.error is None
    assert big.result[3] == h["total"] and big.result[2] == h["time"]
    assert sha(big.result[0]) == h["cu_sha256"] and sha(big.result[1]) == h["gp_sha256"]
