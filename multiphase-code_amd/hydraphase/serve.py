"""Request loop: continuous batching of flow cases on one GPU.

Requests arrive (Poisson) carrying a case: mesh, parameters, initial state
and the number of steps to run.  The loop

* admits waiting requests while a case slot is free and their device
  bytes (hp_case_bytes) fit in the HBM budget: in arrival order (policy
  "fifo") or smallest work first (policy "shortest": active cells x steps,
  ties by arrival), and, with a fill target, only while the resident
  active cells stay within it (a request above the target on its own runs
  when nothing else is resident); with a wait bound (max_wait_s), a
  request that has waited longer than the bound goes ahead of smaller
  work, oldest first, so a stream of small requests cannot hold a large
  one back indefinitely,
* keeps every admitted case resident on the device,
* advances all resident cases together, one hp_batch_step of at most
  `chunk` steps at a time (each case with its own dt), and
* retires a case as soon as it has run its steps: its state is copied back
  and its device memory freed, which lets waiting requests in, and
* after every chunk checks the health of every resident case (hp_batch_health:
  a non-finite state or a non-positive pressure) and retires a failed case
  at once, with its error recorded and no result, so that it holds neither
  a slot nor HBM.

Requests that arrive while others are running join at the next chunk
boundary, so their work overlaps on the device with the cases already
there.  Every resident case advances by a chunk whatever its size, so a
small request shares each step with the whole batch; with a fill target
that just saturates the GPU, smallest-first admission lets short requests
through without holding the GPU below full use.  Shortest-first alone can
hold a large request back while smaller ones keep arriving; the wait bound
ages such a request to the front of the queue.  A case's arithmetic reads
only its own data, so its result is the same bit for bit whether it ran
alone or in any mix.

Reported: throughput in cell-stage updates/s over the makespan, the
time-to-solution (arrival to result on the host) p50/p99 of the requests
that completed, and the number that failed.
"""
from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import batch as HB
from . import mesh as M
from .solver import Solver


@dataclass
class Request:
    rid: int
    spec: HB.CaseSpec
    arrival: float                 # seconds after the start of the run
    mesh: M.Mesh = None
    params: dict = None
    cu: np.ndarray = None
    gp: np.ndarray = None
    nbytes: int = 0
    admit: float = math.nan
    finish: float = math.nan
    result: tuple = None           # (cu, gp, time, total) after the run
    error: str = None              # set when the case failed (no result)
    extra: dict = field(default_factory=dict)


def poisson_requests(n: int, rate: float = 20.0, seed: int = 1, nmin: int = 10, nmax: int = 70,
                     steps=(50, 500), precision: int = 1) -> list[Request]:
    """n requests with exponential inter-arrival times (rate per second);
    cases as the batched ensemble (n log-uniform in [nmin, nmax], Mach,
    radius and viscosity as there) with steps uniform in `steps`.  The
    payload (mesh and state) is built here, before the clock starts: it is
    what a client would send."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    specs = HB.ensemble_specs(n, seed=seed, nmin=nmin, nmax=nmax, steps=steps)
    out = []
    for k, (a, sp) in enumerate(zip(arrivals, specs)):
        m, p, cu, gp, _ = HB.build_case(sp, precision)
        out.append(Request(k, sp, float(a), m, p, cu, gp, HB.case_bytes(m, precision)))
    return out


class Server:
    def __init__(self, budget_bytes: float = 200e9, max_cases: int = 256, chunk: int = 25,
                 graph: bool = False, grid=(0, 0, 0), bytes_per_cell: float = 1000.0,
                 check_health: bool = True, policy: str = "fifo", fill_cells: int | None = None,
                 max_wait_s: float | None = None):
        if policy not in ("fifo", "shortest"):
            raise ValueError(f"unknown admission policy {policy!r} (fifo or shortest)")
        if max_wait_s is not None and not max_wait_s >= 0.0:
            raise ValueError(f"max_wait_s must be a non-negative number of seconds, not {max_wait_s!r}")
        self.policy = policy
        self.fill_cells = None if fill_cells is None else int(fill_cells)
        self.max_wait_s = None if max_wait_s is None else float(max_wait_s)
        self.budget = int(budget_bytes)
        self.chunk = int(chunk)
        self.graph = graph
        self.check_health = check_health
        max_cells = int(self.budget / bytes_per_cell)
        self.batch = HB.Batch(max_cases, max_cells, max(1, max_cells // 4))
        self.batch.set_grid(*grid)
        self.max_cases = max_cases

    def close(self):
        self.batch.close()

    def run(self, requests: list[Request], log=None) -> dict:
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting: collections.deque[Request] = collections.deque()
        active: dict[int, tuple[Request, Solver, int]] = {}
        slot_of: dict[int, int] = {}
        failed = 0
        used = 0
        resident = 0  # active cells of the admitted cases
        peak = 0
        cells_of = lambda r: int(r.mesh.iblank.sum())
        updates = 0
        gpu_ms = 0.0
        for r in requests:
            if r.nbytes > self.budget:
                raise ValueError(f"request {r.rid} needs {r.nbytes} bytes, above the budget")
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0
        while pending or waiting or active:
            now = clock()
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            if self.policy == "shortest":
                # requests past the wait bound first, oldest first; then
                # smallest work, ties by arrival
                aged = (lambda r: self.max_wait_s is not None and now - r.arrival > self.max_wait_s)
                waiting = collections.deque(sorted(
                    waiting, key=lambda r: (0, r.arrival, 0) if aged(r) else (1, cells_of(r) * r.spec.steps, r.arrival)))
            while waiting and len(active) < self.max_cases and used + waiting[0].nbytes <= self.budget:
                if self.fill_cells is not None and active and resident + cells_of(waiting[0]) > self.fill_cells:
                    break
                r = waiting.popleft()
                s = Solver(r.mesh, r.params)
                s.upload(r.cu, r.gp, 0.0, 0)
                slot_of[r.rid] = self.batch.add(s)
                active[r.rid] = (r, s, r.spec.steps)
                used += r.nbytes
                resident += cells_of(r)
                peak = max(peak, used)
                r.admit = clock()
            if not active:
                if pending:
                    time.sleep(max(0.0, pending[0].arrival - clock()))
                continue
            k = min(self.chunk, min(rem for (_, _, rem) in active.values()))
            cells = sum(int(r.mesh.iblank.sum()) for (r, _, _) in active.values())
            self.batch.step(k, graph=self.graph)
            self.batch.synchronize()
            gpu_ms += self.batch.last_ms()
            updates += 4 * k * cells
            health = self.batch.health() if self.check_health else {}
            for rid in list(active):
                r, s, rem = active[rid]
                rem -= k
                bad = health.get(slot_of[rid], 0)
                if rem > 0 and not bad:
                    active[rid] = (r, s, rem)
                    continue
                self.batch.remove(s)
                if bad:
                    r.error = (f"state failed its health check (status {bad}) after "
                               f"{r.spec.steps - rem} of {r.spec.steps} steps")
                    failed += 1
                else:
                    r.result = s.download()
                s.close()
                del active[rid]
                del slot_of[rid]
                used -= r.nbytes
                resident -= cells_of(r)
                r.finish = clock()
                if log:
                    log(f"request {r.rid} (n={r.spec.n}, {r.spec.steps} steps) "
                        f"{'failed' if bad else 'done'} at {r.finish:.2f} s")
        makespan = clock()
        tts = np.array([r.finish - r.arrival for r in requests if r.error is None])
        pct = (lambda q: float(np.percentile(tts, q))) if tts.size else (lambda q: math.nan)
        return {"requests": len(requests), "failed": failed, "makespan_s": makespan, "updates": updates,
                "updates_per_s": updates / makespan, "gpu_busy_s": gpu_ms * 1e-3,
                "tts_p50_s": pct(50), "tts_p99_s": pct(99), "peak_resident_bytes": peak}


def perturbed(cu, rel: float, seed: int = 2024):
    """cu with every entry multiplied by 1 + rel * N(0, 1) (seeded): the
    input noise whose effect on the FP64 step measures the problem's
    conditioning floor (rel = 2**-24: FP32 unit roundoff)."""
    rng = np.random.default_rng(seed)
    return cu * (1.0 + rel * rng.standard_normal(cu.shape))


def alone(r: Request, precision: int | None = None, noise: float = 0.0):
    """The request's case run on its own (hp_step), optionally in another
    precision or from its initial state perturbed by relative noise
    (perturbed): the reference the mix is checked against."""
    p = dict(r.params)
    if precision is not None:
        p["precision"] = precision
    s = Solver(r.mesh, p)
    s.upload(perturbed(r.cu, noise) if noise else r.cu, r.gp, 0.0, 0)
    s.step(r.spec.steps)
    out = s.download()
    s.close()
    return out


def mixture_rel_l2(cu, ref):
    """Relative L2 error of mixture mass, momentum and energy (both phases
    summed), per quantity."""
    mix, mref = cu[:5] + cu[5:], ref[:5] + ref[5:]
    return np.sqrt(((mix - mref) ** 2).sum(axis=1) / (mref ** 2).sum(axis=1))
