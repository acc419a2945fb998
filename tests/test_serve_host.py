"""The request loop's host logic (hydraphase.serve.Server.run) on the CPU,
with stand-in batch and solver objects: admission in arrival order within
the HBM budget and the slot limit, smallest-work-first admission and the
fill target, chunked stepping, retirement when done, retirement of a case
whose health check fails, and the statistics."""
import math
import time
import types

import numpy as np
import pytest

from hydraphase import batch as HB
from hydraphase import serve as S


class FakeSolver:
    live = 0

    def __init__(self, mesh, params):
        self.steps = 0
        self.bad_after = params.get("bad_after")
        FakeSolver.live += 1

    def upload(self, cu, gp, t, tot):
        self.cu = cu

    def download(self):
        return self.cu, None, float(self.steps), self.steps

    def close(self):
        FakeSolver.live -= 1


class FakeBatch:
    instances = []
    step_seconds = 0.0  # wall time a step takes (for the wait-bound test)

    def __init__(self, max_cases, max_cells, max_ghosts):
        self.members = {}
        self.max_cases = max_cases
        self.resident_log = []
        FakeBatch.instances.append(self)

    def set_grid(self, *g):
        pass

    def add(self, s):
        slot = min(set(range(self.max_cases)) - set(self.members))
        self.members[slot] = s
        return slot

    def remove(self, s):
        for k, v in list(self.members.items()):
            if v is s:
                del self.members[k]

    def step(self, n, graph=True):
        self.resident_log.append(len(self.members))
        if FakeBatch.step_seconds:
            time.sleep(FakeBatch.step_seconds * n)
        for s in self.members.values():
            s.steps += n

    def synchronize(self):
        pass

    def last_ms(self):
        return 1.0

    def health(self):
        return {k: (1 if s.bad_after is not None and s.steps >= s.bad_after else 0)
                for k, s in self.members.items()}

    def close(self):
        pass


def requests(nbytes, steps, bad=(), cells=None):
    out = []
    for k, (b, st) in enumerate(zip(nbytes, steps)):
        spec = HB.CaseSpec(4, 0.3, 0.1, False, k, steps=st)
        mesh = types.SimpleNamespace(iblank=np.ones(10 if cells is None else cells[k], dtype=np.int32))
        params = {"bad_after": 3 if k in bad else None}
        out.append(S.Request(k, spec, 0.0, mesh, params, np.full((10, 10), float(k)), None, b))
    return out


def run(monkeypatch, reqs, **kw):
    monkeypatch.setattr(S.HB, "Batch", FakeBatch)
    monkeypatch.setattr(S, "Solver", FakeSolver)
    srv = S.Server(**kw)
    stats = srv.run(reqs)
    srv.close()
    return stats, FakeBatch.instances[-1]


def test_budget_and_slots_respected(monkeypatch):
    reqs = requests([40, 40, 40, 40, 40], [5, 10, 5, 5, 5])
    stats, b = run(monkeypatch, reqs, budget_bytes=100, max_cases=8, chunk=5)
    assert stats["peak_resident_bytes"] <= 100 and max(b.resident_log) == 2
    assert stats["failed"] == 0 and FakeSolver.live == 0
    for r in reqs:
        assert r.error is None and r.result[3] == r.spec.steps  # each ran exactly its steps
    # arrival order: a later request never starts before an earlier one
    admits = [r.admit for r in reqs]
    assert admits == sorted(admits)
    assert stats["updates"] == sum(4 * r.spec.steps * 10 for r in reqs)
    reqs = requests([10] * 6, [4] * 6)
    stats, b = run(monkeypatch, reqs, budget_bytes=1000, max_cases=3, chunk=4)
    assert max(b.resident_log) == 3


def test_failed_case_retired_and_reported(monkeypatch):
    reqs = requests([10, 10, 10], [9, 9, 9], bad=(1,))
    stats, b = run(monkeypatch, reqs, budget_bytes=1000, max_cases=4, chunk=3)
    assert stats["failed"] == 1
    bad = reqs[1]
    assert bad.result is None and "health" in bad.error and "3 of 9" in bad.error
    assert all(r.error is None and r.result[3] == 9 for r in (reqs[0], reqs[2]))
    assert not math.isnan(stats["tts_p50_s"])
    assert FakeSolver.live == 0


def test_health_check_can_be_disabled(monkeypatch):
    reqs = requests([10, 10], [6, 6], bad=(0,))
    stats, _ = run(monkeypatch, reqs, budget_bytes=1000, max_cases=4, chunk=3, check_health=False)
    assert stats["failed"] == 0 and all(r.result[3] == 6 for r in reqs)


def test_shortest_work_first(monkeypatch):
    """All requests wait at the start; one slot: they are served smallest
    active cells x steps first (ties by arrival), and each still runs
    exactly its own steps."""
    cells = [10, 10, 30, 5, 10]
    steps = [8, 2, 4, 4, 2]      # work 80, 20, 120, 20, 20
    reqs = requests([10] * 5, steps, cells=cells)
    stats, b = run(monkeypatch, reqs, budget_bytes=1000, max_cases=1, chunk=2, policy="shortest")
    order = [r.rid for r in sorted(reqs, key=lambda r: r.admit)]
    assert order == [1, 3, 4, 0, 2]
    assert all(r.error is None and r.result[3] == r.spec.steps for r in reqs)
    assert stats["updates"] == sum(4 * c * st for c, st in zip(cells, steps))
    # fifo keeps arrival order
    reqs = requests([10] * 5, steps, cells=cells)
    run(monkeypatch, reqs, budget_bytes=1000, max_cases=1, chunk=2)
    assert [r.rid for r in sorted(reqs, key=lambda r: r.admit)] == [0, 1, 2, 3, 4]


def test_fill_target(monkeypatch):
    """With a fill target the resident active cells stay within it; a
    request larger than the target alone runs when nothing else is
    resident."""
    reqs = requests([10] * 5, [4] * 5, cells=[10, 10, 10, 40, 10])
    stats, b = run(monkeypatch, reqs, budget_bytes=1000, max_cases=8, chunk=2, policy="shortest",
                   fill_cells=25)
    assert max(b.resident_log) == 2
    big = reqs[3]
    for r in reqs:
        if r is not big:
            assert not (r.admit < big.finish and big.admit < r.finish)  # never resident together
    assert all(r.error is None and r.result[3] == 4 for r in reqs) and FakeSolver.live == 0


def test_unknown_policy_rejected():
    with pytest.raises(ValueError):
        S.Server(policy="lifo")


def test_wait_bound_lets_a_large_request_through(monkeypatch):
    """One slot; small requests arrive every 10 ms for 0.4 s and each takes
    four 5 ms steps, so they queue up; a large one arrives at 15 ms.
    Smallest-first alone serves the large one only after every small one;
    with a 50 ms wait bound it goes ahead of the small ones waiting once it
    has waited that long."""
    monkeypatch.setattr(FakeBatch, "step_seconds", 0.005)

    def stream():
        reqs = requests([10] * 41, [4] * 40 + [40], cells=[1] * 40 + [100])
        for k, r in enumerate(reqs):
            r.arrival = 0.01 * k if k < 40 else 0.015
        return reqs

    reqs = stream()
    run(monkeypatch, reqs, budget_bytes=1000, max_cases=1, chunk=1, policy="shortest")
    big = reqs[40]
    assert all(big.admit > r.admit for r in reqs[:40])  # held back behind every small request
    reqs = stream()
    stats, _ = run(monkeypatch, reqs, budget_bytes=1000, max_cases=1, chunk=1, policy="shortest",
                   max_wait_s=0.05)
    big = reqs[40]
    assert 0.065 <= big.admit < 0.25
    assert sum(r.arrival < big.admit < r.admit for r in reqs[:40]) >= 3  # it went ahead of waiting ones
    assert all(r.error is None and r.result[3] == r.spec.steps for r in reqs)
    with pytest.raises(ValueError):
        S.Server(policy="shortest", max_wait_s=-1.0)
