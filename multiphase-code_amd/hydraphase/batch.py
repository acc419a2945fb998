"""Ensembles of independent cases advanced by shared launches (hp_batch).

A Batch holds Solver cases; one hp_batch_step advances every member by the
same number of time steps, each case with its own dt, time and step count.
The kernels walk a device table of 64-cell tiles, so one grid covers all
cases and one captured HIP graph of a step stays valid while cases join and
leave (include/hydra_hip.h).

``tune`` is the autotuner: it times the launch shapes (blocks per launch
family: per-cell kernels, four-lanes-per-cell face kernels, ghost kernels)
and the cell order of the cases (as generated, or Morton-sorted) on a
scratch copy of the ensemble and returns the fastest.  Results are
bitwise independent of those choices and of the other cases in the batch:
every cell's arithmetic reads only its own case.

``ensemble_specs``/``build_case`` generate the uneven ensemble: box meshes of
n^3 hexes x 6 tets, n log-uniform in [nmin, nmax], inflow Mach uniform,
water-sphere radius uniform, gas viscosity on for a given fraction.
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass

import numpy as np

from . import _lib
from . import mesh as M
from ._lib import HydraError
from .solver import Solver

CUS = 256  # compute units of one MI355X


def case_bytes(mesh: M.Mesh, precision: int = 0) -> int:
    """Device bytes a case of this mesh occupies (hp_case_bytes)."""
    return int(_lib.lib().hp_case_bytes(mesh.nodes, mesh.neles, mesh.nghosts, int(precision)))


class Batch:
    def __init__(self, max_cases: int, max_cells: int, max_ghosts: int):
        L = _lib.lib()
        self.L = L
        h = ctypes.c_void_p()
        rc = L.hp_batch_create(ctypes.byref(h), int(max_cases), int(max_cells), int(max_ghosts))
        if rc != 0:
            raise HydraError(f"hp_batch_create failed ({rc}): {L.hp_batch_last_error(None).decode()}")
        self.h = h
        self.max_cases = int(max_cases)
        self.members: dict[int, Solver] = {}
        self.grid = (0, 0, 0)
        self.block = 1

    def _chk(self, rc):
        if rc != 0:
            raise HydraError(f"batch call failed ({rc}): {self.L.hp_batch_last_error(self.h).decode()}")

    def add(self, s: Solver) -> int:
        slot = ctypes.c_int()
        self._chk(self.L.hp_batch_add(self.h, s.h, ctypes.byref(slot)))
        self.members[slot.value] = s
        return slot.value

    def remove(self, s: Solver):
        self._chk(self.L.hp_batch_remove(self.h, s.h))
        for k, v in list(self.members.items()):
            if v is s:
                del self.members[k]

    def set_grid(self, cell: int, quad: int, ghost: int):
        self._chk(self.L.hp_batch_set_grid(self.h, int(cell), int(quad), int(ghost)))
        self.grid = (cell, quad, ghost)

    def set_block(self, tiles_per_block: int):
        """Tiles of 64 cells per block (1, 2 or 4) for the streaming per-cell
        kernels (hp_batch_set_block)."""
        self._chk(self.L.hp_batch_set_block(self.h, int(tiles_per_block)))
        self.block = int(tiles_per_block)

    def step(self, n: int = 1, graph: bool = True):
        self._chk(self.L.hp_batch_step(self.h, int(n), 1 if graph else 0))

    def synchronize(self):
        self._chk(self.L.hp_batch_synchronize(self.h))

    def last_ms(self) -> float:
        v = ctypes.c_double()
        self._chk(self.L.hp_batch_last_ms(self.h, ctypes.byref(v)))
        return v.value

    def health(self) -> dict[int, int]:
        """hp_health status of every member, by slot (hp_batch_health)."""
        n = self.max_cases
        buf = (ctypes.c_int * n)()
        self._chk(self.L.hp_batch_health(self.h, buf, n))
        return {slot: buf[slot] for slot in self.members}

    def close(self):
        if getattr(self, "h", None):
            self.L.hp_batch_destroy(self.h)
            self.h = None
            self.members = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------------------
# the uneven ensemble
# --------------------------------------------------------------------------

@dataclass
class CaseSpec:
    n: int          # hexes per edge; n^3 x 6 tets
    mach: float     # inflow Mach number (gas sound speed)
    radius: float   # water sphere radius (unit box)
    viscous: bool   # gas viscosity on
    seed: int       # mesh jitter seed
    steps: int = 200


def ensemble_specs(n_cases: int = 64, seed: int = 1, nmin: int = 8, nmax: int = 48,
                   mach=(0.1, 0.8), radius=(0.05, 0.25), viscous_fraction: float = 0.5,
                   steps=(200, 200)) -> list[CaseSpec]:
    """Cases with n log-uniform in [nmin, nmax], Mach and radius uniform,
    gas viscosity on for viscous_fraction of them, steps uniform (inclusive)."""
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(math.log(nmin), math.log(nmax + 1), n_cases)).astype(int).clip(nmin, nmax)
    ma = rng.uniform(mach[0], mach[1], n_cases)
    ra = rng.uniform(radius[0], radius[1], n_cases)
    visc = np.zeros(n_cases, dtype=bool)
    visc[rng.permutation(n_cases)[: int(round(viscous_fraction * n_cases))]] = True
    st = rng.integers(steps[0], steps[1] + 1, n_cases)
    return [CaseSpec(int(n[k]), float(ma[k]), float(ra[k]), bool(visc[k]), seed * 1000 + k, int(st[k]))
            for k in range(n_cases)]


def build_case(spec: CaseSpec, precision: int = 0, morton: bool = False, myflux: int = 2,
               slabs: bool = False):
    """(mesh, params, cu, gp, perm) of one ensemble member.  morton: cells
    sorted along a Z-order curve; slabs: cells numbered in y-z slabs of 8 x
    8 hex layers (mesh.slab_order).  New cell j is generated cell perm[j];
    the results per cell are the same, only the memory order changes; perm
    is None for the generated order."""
    m = M.box_mesh(spec.n, spec.n, spec.n, jitter=0.1, seed=spec.seed)
    p = M.water_air_params(mach=spec.mach, cfl=0.5, viscous_gas=spec.viscous, myflux=myflux)
    p["precision"] = precision
    cu, gp = M.sphere_state(m, p, [(0.0, 0.0, 0.0, spec.radius)])
    perm = None
    if morton or slabs:
        perm = M.morton_order(m) if morton else M.slab_order(m, spec.n, spec.n)
        m = m.permuted(perm)
        cu = np.ascontiguousarray(cu[:, perm])  # ghost order (and gp) is kept
    return m, p, cu, gp, perm


def capacity(meshes) -> tuple[int, int, int]:
    return len(meshes), sum(m.neles for m in meshes), sum(m.nghosts for m in meshes)


# --------------------------------------------------------------------------
# autotuner
# --------------------------------------------------------------------------

GRID_CHOICES = [0, 2 * CUS, 4 * CUS, 8 * CUS, 16 * CUS]
BLOCK_CHOICES = [1, 2, 4]


def _time_steps(b: Batch, steps: int) -> float:
    b.step(1)           # capture (if needed) and warm
    b.synchronize()
    b.step(steps)
    return b.last_ms() / steps


def tune(specs, precision: int = 1, steps: int = 3, myflux: int = 2, log=None) -> dict:
    """Pick the launch shape and cell order for this ensemble.

    Builds a scratch copy of the ensemble (so the real cases do not move),
    then walks the choices one family at a time (coordinate descent): cell
    order, then cell-kernel blocks, face-kernel blocks and ghost-kernel
    blocks, each from GRID_CHOICES (0 = one block per tile), then the block
    size of the streaming per-cell kernels (BLOCK_CHOICES tiles of 64 cells
    per block).  The cell orders tried: as generated, Z order (morton) and
    y-z slabs (slabs).  Returns {"morton", "slabs", "grid": (cell, quad,
    ghost), "block", "ms_per_step", "trials"}."""
    trials = []

    def build(order):
        cases = [build_case(s, precision, order["morton"], myflux, order["slabs"]) for s in specs]
        ms = [c[0] for c in cases]
        b = Batch(*capacity(ms))
        sols = []
        for (m, p, cu, gp, _) in cases:
            s = Solver(m, p)
            s.upload(cu, gp, 0.0, 0)
            b.add(s)
            sols.append(s)
        return b, sols

    best = None
    for order in ({"morton": False, "slabs": False}, {"morton": True, "slabs": False},
                  {"morton": False, "slabs": True}):
        b, sols = build(order)
        t = _time_steps(b, steps)
        trials.append((dict(order, grid=(0, 0, 0)), t))
        if log:
            log(f"tune: order={order} grid=(0,0,0) {t:.3f} ms/step")
        if best is None or t < best[1]:
            best = (dict(order, grid=(0, 0, 0)), t)
        b.close()
        for s in sols:
            s.close()
    order = {"morton": best[0]["morton"], "slabs": best[0]["slabs"]}
    b, sols = build(order)
    grid = list(best[0]["grid"])
    best_t = best[1]
    for fam in range(3):
        for g in GRID_CHOICES:
            if g == grid[fam]:
                continue
            trial = list(grid)
            trial[fam] = g
            b.set_grid(*trial)
            t = _time_steps(b, steps)
            trials.append((dict(order, grid=tuple(trial)), t))
            if log:
                log(f"tune: grid={tuple(trial)} {t:.3f} ms/step")
            if t < best_t:
                best_t, grid = t, trial
    b.set_grid(*grid)
    block = 1
    for tpb in BLOCK_CHOICES[1:]:
        b.set_block(tpb)
        t = _time_steps(b, steps)
        trials.append((dict(order, grid=tuple(grid), block=tpb), t))
        if log:
            log(f"tune: grid={tuple(grid)} block={tpb} {t:.3f} ms/step")
        if t < best_t:
            best_t, block = t, tpb
    b.close()
    for s in sols:
        s.close()
    return dict(order, grid=tuple(grid), block=block, ms_per_step=best_t, trials=trials)


def run_ensemble(specs, precision: int = 1, tuning: dict | None = None, graph: bool = True,
                 myflux: int = 2):
    """Build the ensemble, advance every case by its own step count in one
    batch (cases leave when done), and return per-case final states, their
    hp_health status when they left, and the timing: {"ms": total,
    "updates": cell-stage updates, "states": [...], "health": [...]}."""
    tuning = tuning or {"morton": False, "grid": (0, 0, 0)}
    cases = [build_case(s, precision, tuning["morton"], myflux, tuning.get("slabs", False)) for s in specs]
    b = Batch(*capacity([c[0] for c in cases]))
    b.set_grid(*tuning["grid"])
    b.set_block(tuning.get("block", 1))
    sols, slots = [], []
    for (m, p, cu, gp, _) in cases:
        s = Solver(m, p)
        s.upload(cu, gp, 0.0, 0)
        slots.append(b.add(s))
        sols.append(s)
    remaining = [s.steps for s in specs]
    active = [int(c[0].iblank.sum()) for c in cases]
    updates, total_ms, done = 0, 0.0, 0
    states = [None] * len(specs)
    health = [0] * len(specs)
    t0 = time.perf_counter()
    while any(r > 0 for r in remaining):
        k = min(r for r in remaining if r > 0)
        b.step(k, graph=graph)
        total_ms += b.last_ms()
        updates += 4 * k * sum(int(a) for a, r in zip(active, remaining) if r > 0)
        status = None
        for i, r in enumerate(remaining):
            if r <= 0:
                continue
            remaining[i] = r - k
            if remaining[i] == 0:
                if status is None:
                    status = b.health()  # outside the timed steps
                health[i] = status[slots[i]]
                b.remove(sols[i])
                states[i] = sols[i].download()
                done += 1
    wall = time.perf_counter() - t0
    b.close()
    perms = [c[4] for c in cases]
    for s in sols:
        s.close()
    return {"ms": total_ms, "wall_s": wall, "updates": updates, "states": states, "health": health,
            "meshes": [c[0] for c in cases], "perms": perms}
