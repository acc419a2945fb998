"""Throughput of the HydraPhase per-step FV update on one MI355X.

Workload (the single large case, two overlapping meshes): a 96 x 96 x 90
hex box (x, y, z in [-0.5, 0.5]) Kuhn-split into tets with interior nodes
jittered by +-10% of the spacing (seed 1), its tets of centroid radius
< 0.15 from the z axis cut out, plus a near-body annulus around the
cylinder r = 0.1 (no-slip wall, type 5 ghosts) out to r = 0.25, 10 x 112 x
90 curved hexes x 6 = 604,800 tets; the two are coupled by fixed
barycentric donor lists (two fringe layers each side, applied after every
RK stage).  Gas viscosity on (air 1.8e-5 Pa s via Sutherland, water 1e-3
Pa s), a train of 8 water spheres (radii 0.03-0.08, seeded) in an air
stream at Mach 0.3, CFL 0.5.  FP64 throughout (--precision 1: FP32 bulk).
--single-mesh runs the box alone with the cylinder as a wall hole.  The
box's cells are numbered in y-z slabs of 8 x 8 hex layers (x fastest within
a slab; hydraphase.mesh.slab_order) so that face neighbours sit close in
memory; per-cell results do not depend on the numbering.

A step is one full time step of the reference (main.F90:515-625): primitive,
CFL reduction, JST smoothing, viscous fluxes and four RK stages of
primitiveplusbc + gradients + limiter + HLLC face flux + source + update
(+ donor gather).  The metric counts cell-stage updates (active cells x 4
stages) per second with the state resident in HBM.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n NX NY NZ]
       python bench.py summarize-counters READREQ.csv WRITEREQ.csv CALIB_MB
       python bench.py summarize-trace results.db [KERNEL-FOR-A-PER-GRID-SPLIT]
       python bench.py summarize-instructions counter_collection.csv
       python bench.py summarize-pmc counter_collection.csv
       python bench.py summarize-gaps results.db [WINDOW-US]
       python bench.py run-ensemble [graph|eager] [STEPS]

roofline.traffic is the HBM bytes of one k_flux launch measured with two
rocprofv3 --pmc runs, of the L2's read requests by size
(TCC_EA0_RDREQ_32B/_64B/_128B, the terms of FETCH_SIZE) and of its write
requests (TCC_EA0_WRREQ/_64B, those of WRITE_SIZE), each run with
--calib-copy-mb 1024 (the library's streaming copy kernel, of known traffic)
and summarised by summarize-counters into
profiles/k_flux_overset96x96x90_fp64_counters.md.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "multiphase-code_amd"))

HBM_PEAK_GBS = 8000.0
FLUX_TIMING_STEPS = 5
# algorithmic bytes of one face-flux launch per active cell (FP64): own
# primitives 14x8, gradients 33x8, limiter 8, area vectors 12x8, face
# centroids 12x8, centroid 3x8, neighbour ids 4x4, vol 8, interphase source
# terms 5x8, viscous flux 8x8, d24 10x8, cu 10x8, cn read (stages 2-4)
# 3/4x10x8, cn1 read (stage 4) and write (stage 1) 2x1/4x10x8, cn write
# 10x8; and the fused primitive recovery of the updated cells: the next
# stage's primitives 14x8 and, at the last stage, cu 1/4x10x8.  Neighbour
# values are counted once, by their owning cell.
FLUX_BYTES_PER_CELL = 112 + 264 + 8 + 96 + 96 + 24 + 16 + 8 + 40 + 64 + 80 + 80 + 60 + 40 + 80 + 112 + 20
# FP32-bulk form: gradients 33x4 and limiter 4 instead of 33x8 and 8; the
# source is formed in the flux from vol+dl (16) instead of read (8 + 40)
FLUX_BYTES_PER_CELL_F32 = FLUX_BYTES_PER_CELL - 132 - 4 + 16 - 48
# The same launch priced by the path's own counting rule, independent of
# this library's tables: the face-flux pass reads primitives 104, gradients
# 264, limiter 8, area vectors 96, centroid 24, node ids 16, node
# coordinates 4 (a node per six cells), neighbour ids 16, vol 8, dl 8,
# viscous flux 64, d24 80, cu 80, cn 80 and cn1 20 (stage 4, a quarter),
# writes cn 80 and cn1 20 (stage 1, a quarter): 972 B; the primitive
# recovery fused into it reads cn 80 and writes 13 primitives 104: 184 B.
# roofline.achieved and .frac use these 1,156 B per active cell.
FLUX_BYTES_RULE = 972 + 184
# FP32 bulk under the same rule: gradients 132 and limiter 4
FLUX_BYTES_RULE_F32 = FLUX_BYTES_RULE - 132 - 4


def build_case(nx, ny, nz, seed=1, precision=0, overset=True, slabs=(8, 8)):
    """(mesh, params, cu, gp, donors) of the single large case; donors is
    (recv, ids, weights) or None for the single-mesh form.  The box's cells
    are numbered in y-z slabs of `slabs` (hydraphase.mesh.slab_order) for
    gather locality (None: as generated); per-cell results do not depend
    on the numbering."""
    from hydraphase import mesh as M
    from hydraphase import overset as OV
    hole = (0.0, 0.0, 0.1)
    donors = None
    if overset:
        m, recv, ids, w = OV.overset_mesh(nx, ny, nz, nr=10, ntheta=112, r_body=0.1, r_outer=0.25,
                                          r_cut=0.15, fringe=2, jitter=0.1, seed=seed, slabs=slabs)
        donors = (recv, ids, w)
    else:
        m = M.box_mesh(nx, ny, nz, jitter=0.1, seed=seed, hole=hole)
        if slabs is not None:
            m = m.permuted(M.slab_order(m, ny, nz, by=slabs[0], bz=slabs[1]))
    p = M.water_air_params(cfl=0.5, viscous_gas=True)
    p["precision"] = precision
    spheres = M.sphere_train(seed, n=8, rmin=0.03, rmax=0.08, avoid=(0.0, 0.0, 0.25 if overset else 0.1))
    cu, gp = M.sphere_state(m, p, spheres)
    return m, p, cu, gp, donors


def stream_copy_gbs(mib=1024, reps=10):
    """Measured copy ceiling: the library's hand-written streaming copy kernel
    (hp_stream_bandwidth) copies mib MiB reps times; GB/s of read + write
    bytes."""
    import ctypes
    from hydraphase import _lib
    L = _lib.lib()
    gbs, ms = ctypes.c_double(), ctypes.c_double()
    rc = L.hp_stream_bandwidth(mib << 20, reps, ctypes.byref(gbs), ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError(f"hp_stream_bandwidth failed: {L.hp_last_error(None).decode()}")
    return gbs.value


def cpu_baseline(m, p, cu, gp, donors=None, seconds_target=20.0):
    """The serial C oracle (a port of the reference's FP64 path, 1 thread)
    on a bounded sample of the same workload: whole steps of the same case
    until about seconds_target of CPU time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    o = O.Oracle(m, p)
    o.set_state(cu, gp, 0.0, 0)
    if donors is not None:
        o.set_donors(*donors)
        o.call("donor_gather")
    nsteps, t0 = 0, time.perf_counter()
    while True:
        o.step()
        nsteps += 1
        el = time.perf_counter() - t0
        if el >= seconds_target or nsteps >= 50:
            break
    active = int((m.iblank == 1).sum())
    return {"value": active * 4 * nsteps / el, "unit": "cell-stage updates/s", "cores": 1,
            "kind": "port",
            "sample": f"{nsteps} full step(s) of this workload ({active} cells) on the C oracle, "
                      f"{el:.1f} s"}


PROFILES = os.path.join(ROOT, "profiles")
TRAFFIC_FILE = os.path.join(PROFILES, "k_flux_overset96x96x90_fp64_counters.md")


def flux_traffic_bytes():
    """HBM bytes per k_flux launch at this workload, as measured by the
    FETCH_SIZE/WRITE_SIZE counter runs summarised in TRAFFIC_FILE (None if
    that summary is absent)."""
    import re
    try:
        txt = open(TRAFFIC_FILE).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch k_flux: (\d+)", txt)
    return int(m.group(1)) if m else None


def _pmc_means(path):
    """{kernel: {counter: mean per dispatch}} and {kernel: dispatches} of a
    rocprofv3 --pmc counter_collection.csv."""
    import csv
    import collections
    tot = collections.defaultdict(lambda: collections.defaultdict(float))
    disp = collections.defaultdict(set)
    with open(path) as fh:
        for r in csv.DictReader(fh):
            tot[r["Kernel_Name"]][r["Counter_Name"]] += float(r["Counter_Value"])
            disp[r["Kernel_Name"]].add(r["Dispatch_Id"])
    return ({k: {c: v / len(disp[k]) for c, v in cs.items()} for k, cs in tot.items()},
            {k: len(v) for k, v in disp.items()})


def summarize_counters(read_csv, write_csv, calib_mb):
    """HBM bytes per dispatch from the L2's memory requests by size: a run of
    TCC_EA0_RDREQ_32B/_64B/_128B (and the total TCC_EA0_RDREQ) and a run of
    TCC_EA0_WRREQ/_64B.  Read bytes = 32 n32 + 64 n64 + 128 n128, written =
    32 (n - n64) + 64 n64.  The calibration kernel (hydra::k_stream_copy,
    reads and writes calib_mb MiB) checks the accounting; its ratio of known
    to counted bytes is printed and applied.  Prints markdown."""
    rd, nrd = _pmc_means(read_csv)
    wr, _ = _pmc_means(write_csv)

    def rbytes(c):
        return (32 * c.get("TCC_EA0_RDREQ_32B_sum", 0.0) + 64 * c.get("TCC_EA0_RDREQ_64B_sum", 0.0)
                + 128 * c.get("TCC_EA0_RDREQ_128B_sum", 0.0))

    def wbytes(c):
        n, n64 = c.get("TCC_EA0_WRREQ_sum", 0.0), c.get("TCC_EA0_WRREQ_64B_sum", 0.0)
        return 32 * (n - n64) + 64 * n64

    cal = [k for k in rd if "k_stream_copy" in k]
    if not cal:
        raise SystemExit("calibration kernel hydra::k_stream_copy not found in " + read_csv)
    known = calib_mb * float(1 << 20)
    rfac = known / rbytes(rd[cal[0]])
    wfac = known / wbytes(wr[cal[0]])
    c = rd[cal[0]]
    print(f"calibration kernel `hydra::k_stream_copy`: reads and writes {calib_mb} MiB; counted "
          f"{rbytes(c) / known:.4f} x read ({c.get('TCC_EA0_RDREQ_128B_sum', 0):.0f} x 128 B, "
          f"{c.get('TCC_EA0_RDREQ_64B_sum', 0):.0f} x 64 B, {c.get('TCC_EA0_RDREQ_32B_sum', 0):.0f} x 32 B "
          f"requests) and {wbytes(wr[cal[0]]) / known:.4f} x written -> factors {rfac:.4f} (read), "
          f"{wfac:.4f} (write)\n")
    print("| kernel | dispatches | read req 128/64/32 B (share of requests) | HBM read MB | HBM write MB |")
    print("|---|---|---|---|---|")
    flux = face32 = None
    for k in rd:
        if "hydra" not in k or "k_stream_copy" in k:
            continue
        c = rd[k]
        n = max(c.get("TCC_EA0_RDREQ_sum", 0.0), 1.0)
        mix = "/".join(f"{c.get(f'TCC_EA0_RDREQ_{b}B_sum', 0.0) / n:.2f}" for b in (128, 64, 32))
        r = rbytes(c) * rfac / 1e6
        w = wbytes(wr.get(k, {})) * wfac / 1e6
        name = k.split("(")[0].replace("void ", "")
        print(f"| `{name}` | {nrd[k]} | {mix} | {r:.1f} | {w:.1f} |")
        if name.startswith("hydra::k_flux"):
            flux = int(round((r + w) * 1e6))
        if name.startswith("hydra::k_face32<") or name.startswith("hydra::k_cells32<"):
            face32 = (face32 or 0) + int(round((r + w) * 1e6))
    if flux is not None:
        print(f"\ntraffic_bytes_per_launch k_flux: {flux}")
    if face32 is not None:
        print(f"\ntraffic_bytes_per_stage k_face32 + k_cells32: {face32}")


def summarize_pmc(path, kernels=("k_flux", "k_face32<", "k_cells32", "k_grad", "k_viscous", "k_stream_copy")):
    """Mean per dispatch of every counter of one rocprofv3 --pmc run, for the
    kernels named (substrings).  Prints markdown."""
    means, nd = _pmc_means(path)
    names = sorted({c for cs in means.values() for c in cs})
    print("| kernel | dispatches | " + " | ".join(names) + " |")
    print("|---|---|" + "---|" * len(names))
    for k, cs in means.items():
        if not any(s in k for s in kernels):
            continue
        name = k.split("(")[0].replace("void ", "")
        print(f"| `{name}` | {nd[k]} | " + " | ".join(f"{cs.get(c, 0.0):.4g}" for c in names) + " |")


def summarize_instructions(sq_csv):
    """Per-kernel means per dispatch of one rocprofv3 --pmc run of SQ_WAVES,
    SQ_INSTS_VALU, SQ_INSTS_VMEM_RD, SQ_INSTS_LDS, SQ_ACTIVE_INST_VALU and
    SQ_WAVE_CYCLES.  Prints markdown."""
    import csv
    import collections
    tot = collections.defaultdict(lambda: collections.defaultdict(float))
    disp = collections.defaultdict(set)
    with open(sq_csv) as fh:
        for r in csv.DictReader(fh):
            if "hydra" in r["Kernel_Name"]:
                tot[r["Kernel_Name"]][r["Counter_Name"]] += float(r["Counter_Value"])
                disp[r["Kernel_Name"]].add(r["Dispatch_Id"])
    print("| kernel | dispatches | waves | VALU insts | VMEM read insts | LDS insts | VALU active / wave cycles |")
    print("|---|---|---|---|---|---|---|")
    for k, c in tot.items():
        n = len(disp[k])
        name = k.split("(")[0].replace("void ", "")
        act = c["SQ_ACTIVE_INST_VALU"] / max(c["SQ_WAVE_CYCLES"], 1.0)
        print(f"| `{name}` | {n} | {c['SQ_WAVES'] / n:.0f} | {c['SQ_INSTS_VALU'] / n:.3e} | "
              f"{c['SQ_INSTS_VMEM_RD'] / n:.3e} | {c['SQ_INSTS_LDS'] / n:.3e} | {act:.2f} |")


def summarize_trace(db_path, top=20, by_grid=None):
    """Per-kernel totals from a rocprofv3 --kernel-trace --stats database;
    with by_grid (a kernel name substring), that kernel's launches split by
    grid size as well (a run that launches it for cases of several sizes)."""
    import sqlite3
    cur = sqlite3.connect(db_path).cursor()
    print("| kernel | calls | total us | avg us | % |\n|---|---|---|---|---|")
    for r in cur.execute(f"select * from top_kernels limit {int(top)}"):
        name = r[0].replace("|", "/")
        print(f"| `{name}` | {r[1]} | {r[2]:.1f} | {r[3]:.1f} | {r[4]:.2f} |")
    if by_grid:
        print(f"\n`{by_grid}` by grid size (work-items), largest first:\n")
        print("| grid | calls | avg us |\n|---|---|---|")
        q = ("select grid_x, count(*), avg(end - start) from kernels where name like ? "
             "group by grid_x order by grid_x desc limit 8")
        for g, n, d in cur.execute(q, (f"%{by_grid}%",)):
            print(f"| {g} | {n} | {d / 1e3:.1f} |")


def summarize_gaps(db_path, window_us=100.0, min_launches=100):
    """Idle time of the GPU between consecutive kernels of a rocprofv3
    kernel trace: kernels in start order, the idle time before each one its
    start minus the latest end so far.  An idle stretch longer than
    window_us separates bursts (host work between calls); inside the bursts
    of at least min_launches launches (the stepping), the idle time is what
    fusing the step's launches into one persistent launch could save at
    most.  Prints markdown."""
    import sqlite3
    cur = sqlite3.connect(db_path).cursor()
    rows = sorted(cur.execute("select start, end from kernels"))
    st = np.array([r[0] for r in rows], dtype=np.float64)
    en = np.maximum.accumulate(np.array([r[1] for r in rows], dtype=np.float64))
    idle = np.concatenate([[np.inf], (st[1:] - en[:-1]) / 1e3])  # us before each launch
    cuts = np.flatnonzero(idle > window_us).tolist() + [len(rows)]
    print("| burst | launches | span us | kernel busy us | idle us | idle % | median gap us |")
    print("|---|---|---|---|---|---|---|")
    tot_span = tot_idle = 0.0
    for a, b in zip(cuts, cuts[1:]):
        if b - a < min_launches:
            continue
        gaps = np.clip(idle[a + 1:b], 0.0, None)
        span = (en[b - 1] - st[a]) / 1e3
        busy = span - gaps.sum()
        tot_span += span
        tot_idle += gaps.sum()
        print(f"| {a} | {b - a} | {span:.1f} | {busy:.1f} | {gaps.sum():.1f} | "
              f"{100 * gaps.sum() / span:.3f} | {np.median(gaps):.2f} |")
    print(f"\nbursts of >= {min_launches} launches: {tot_idle:.1f} us idle in {tot_span:.1f} us "
          f"({100 * tot_idle / max(tot_span, 1e-9):.3f}%)")


def ensemble_bench(log, steps=200, sample=3, baseline_steps=20):
    """The 64-case uneven ensemble in its FP32-bulk batched form (autotuned
    launch shape, graph-captured step) against the plain one-case FP64 path
    (each case alone through hp_step) on the same cases; the error of the
    batched results against the FP64 path (bitwise the oracle's) for a
    sample of cases."""
    from hydraphase import Solver
    from hydraphase import batch as HB
    from hydraphase import serve as SV
    specs = HB.ensemble_specs(64, seed=1, steps=(steps, steps))
    t0 = time.time()
    tuning = HB.tune(specs, precision=1, log=log)
    log(f"# tuned in {time.time() - t0:.1f} s: morton={tuning['morton']} slabs={tuning['slabs']} "
        f"grid={tuning['grid']} block={tuning['block']}")
    res = HB.run_ensemble(specs, precision=1, tuning=tuning, graph=True)
    batched = res["updates"] / (res["ms"] * 1e-3)
    # plain one-case FP64 path: each case alone (in the same cell order),
    # baseline_steps steps after 2 warm-up steps
    upd, ms = 0, 0.0
    for sp in specs:
        m, p, cu, gp, _ = HB.build_case(sp, 0, tuning["morton"], 2, tuning["slabs"])
        s = Solver(m, p)
        s.upload(cu, gp, 0.0, 0)
        s.step(2)
        s.step(baseline_steps)
        tot, _, _ = s.last_timing()
        ms += tot
        upd += 4 * baseline_steps * int(m.iblank.sum())
        s.close()
    plain = upd / (ms * 1e-3)
    errs, floors = [], []
    for k in np.linspace(0, len(specs) - 1, sample).astype(int):
        m, p, cu, gp, _ = HB.build_case(specs[k], 0)
        fp64 = []
        for start in (cu, SV.perturbed(cu, 2.0 ** -24)):
            s = Solver(m, p)
            s.upload(start, gp, 0.0, 0)
            s.step(steps)
            fp64.append(s.download()[0])
            s.close()
        ref = fp64[0]
        got = res["states"][k][0]
        if res["perms"][k] is not None:
            un = np.empty_like(got)
            un[:, res["perms"][k]] = got
            got = un
        errs.append(float(SV.mixture_rel_l2(got, ref).max()))
        floors.append(float(SV.mixture_rel_l2(fp64[1], ref).max()))
    cells = sum(int(mm.iblank.sum()) for mm in res["meshes"])
    return {"workload": f"64 boxes n^3 x 6 tets, n log-uniform 8..48 ({cells} tets), Mach 0.1-0.8, "
                        f"R 0.05-0.25, half viscous, {steps} steps each",
            "updates_per_s": batched, "ms_per_step": res["ms"] / steps,
            "plain_fp64_one_case_updates_per_s": plain, "speedup_vs_plain_fp64": batched / plain,
            "tuning": {"morton": tuning["morton"], "slabs": tuning["slabs"], "grid": list(tuning["grid"]),
                       "block": tuning["block"]},
            "mixture_rel_l2_vs_fp64_sample": errs,
            "mixture_rel_l2_floor_sample": floors, "sample_cases": sample,
            "cases_unhealthy_at_end": int(sum(1 for h in res["health"] if h))}


def run_ensemble_steps(mode="graph", steps=10, warmup=2):
    """The 64-case FP32-bulk ensemble of ensemble_bench in one batch with the
    default launch shape (one block per tile), stepped warmup then `steps`
    times, by graph replays or eager launches: the run whose kernel trace
    shows the launch gaps of the batched step.  Prints ms/step."""
    if mode not in ("graph", "eager"):
        raise SystemExit(f"run-ensemble: mode {mode!r} is neither graph nor eager")
    from hydraphase import Solver
    from hydraphase import batch as HB
    specs = HB.ensemble_specs(64, seed=1)
    cases = [HB.build_case(sp, 1) for sp in specs]
    b = HB.Batch(*HB.capacity([c[0] for c in cases]))
    sols = []
    for (m, p, cu, gp, _) in cases:
        s = Solver(m, p)
        s.upload(cu, gp, 0.0, 0)
        b.add(s)
        sols.append(s)
    print(f"# {len(cases)} cases built", file=sys.stderr, flush=True)
    graph = mode == "graph"
    b.step(warmup, graph=graph)
    b.synchronize()
    b.step(steps, graph=graph)
    b.synchronize()
    print(json.dumps({"mode": mode, "steps": steps, "ms_per_step": b.last_ms() / steps}), flush=True)
    b.close()
    for s in sols:
        s.close()


SERVE_MAX_WAIT_S = 4.0


def serve_bench(log, n_requests=48, rate=20.0, sample_frac=0.05, max_wait_s=SERVE_MAX_WAIT_S, seed=7):
    """Poisson arrivals of cases (n 10..70, 50..500 steps), admitted smallest
    work first within a 200 GB HBM budget while at most 1M active cells are
    resident (a request waiting longer than max_wait_s goes first), and
    stepped together (FP32 bulk).  sample_frac = 0: no error sample."""
    from hydraphase import serve as SV
    reqs = SV.poisson_requests(n_requests, rate=rate, seed=seed, nmin=10, nmax=70, steps=(50, 500), precision=1)
    srv = SV.Server(budget_bytes=200e9, max_cases=256, chunk=25, policy="shortest", fill_cells=1_000_000,
                    max_wait_s=max_wait_s)
    try:
        st = srv.run(reqs)
    finally:
        srv.close()
    k = max(1, int(round(sample_frac * n_requests))) if sample_frac > 0 else 0
    done = [r for r in reqs if r.error is None]  # requests retired as failed carry no result
    errs, floors = [], []
    for r in (done[:: max(1, len(done) // k)][:k] if k else []):
        ref = SV.alone(r, precision=0)
        errs.append(float(SV.mixture_rel_l2(r.result[0], ref[0]).max()))
        floors.append(float(SV.mixture_rel_l2(SV.alone(r, precision=0, noise=2.0 ** -24)[0], ref[0]).max()))
    cells = sum(int(r.mesh.iblank.sum()) for r in reqs)
    return {"workload": f"{n_requests} requests, Poisson {rate}/s, boxes n 10..70 ({cells} tets in all), "
                        "50..500 steps, 200 GB budget, smallest work first within 1M resident active cells, "
                        f"requests waiting over {max_wait_s:g} s first, FP32 bulk",
            "updates_per_s": st["updates_per_s"], "tts_p50_s": st["tts_p50_s"], "tts_p99_s": st["tts_p99_s"],
            "makespan_s": st["makespan_s"], "gpu_busy_s": st["gpu_busy_s"],
            "peak_resident_gb": st["peak_resident_bytes"] / 1e9, "requests_failed": st["failed"],
            "mixture_rel_l2_vs_fp64_sample": errs, "mixture_rel_l2_floor_sample": floors}


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "summarize-counters":
        summarize_counters(sys.argv[2], sys.argv[3], int(sys.argv[4]))
        return
    if len(sys.argv) > 1 and sys.argv[1] == "summarize-gaps":
        summarize_gaps(sys.argv[2], *map(float, sys.argv[3:4]))
        return
    if len(sys.argv) > 1 and sys.argv[1] == "summarize-pmc":
        summarize_pmc(sys.argv[2])
        return
    if len(sys.argv) > 1 and sys.argv[1] == "summarize-instructions":
        summarize_instructions(sys.argv[2])
        return
    if len(sys.argv) > 1 and sys.argv[1] == "run-ensemble":
        run_ensemble_steps(*(sys.argv[2:3]), *map(int, sys.argv[3:4]))
        return
    if len(sys.argv) > 1 and sys.argv[1] == "summarize-trace":
        summarize_trace(sys.argv[2], by_grid=sys.argv[3] if len(sys.argv) > 3 else None)
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, nargs=3, default=[96, 96, 90])
    ap.add_argument("--precision", type=int, default=0, choices=[0, 1],
                    help="0: FP64 drop-in (the headline); 1: FP32-bulk variant")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--single-mesh", action="store_true",
                    help="the box alone with the cylinder as a wall hole (no annulus, no donors)")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched 64-case ensemble")
    ap.add_argument("--no-serve", action="store_true", help="skip the request-loop run")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="first run the streaming copy kernel over this many MiB, the kernel of "
                         "known traffic that calibrates the HBM request counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every case runs "
              "on one device; the multi-GPU decomposition is a design only)", file=sys.stderr)
        sys.exit(2)

    if args.calib_copy_mb:
        # hydra::k_stream_copy: one kernel of known traffic (reads and writes
        # exactly calib_copy_mb MiB per launch), the yardstick for the counters
        stream_copy_gbs(args.calib_copy_mb, 1)
    t0 = time.time()
    m, p, cu, gp, donors = build_case(*args.n, precision=args.precision, overset=not args.single_mesh)
    print(f"# case built: {m.neles} tets, {m.nghosts} ghosts, "
          f"{0 if donors is None else donors[0].size} receptors ({time.time() - t0:.1f} s)", file=sys.stderr,
          flush=True)
    from hydraphase import Solver
    s = Solver(m, p)
    s.upload(cu, gp, 0.0, 0)
    if donors is not None:
        s.set_donors(*donors)
        s.apply_donors()
    if args.warmup:
        s.step(args.warmup)
        s.synchronize()
    # the timed steps: hp_step replays one captured step (HIP graph)
    s.step(args.steps)
    s.synchronize()
    total_ms, _, _ = s.last_timing()
    # after them, a few eager steps with HIP events around every flux
    # launch for the roofline figures
    s.set_flux_timing(True)
    s.step(FLUX_TIMING_STEPS)
    s.synchronize()
    _, flux_ms, nflux = s.last_timing()
    s.set_flux_timing(False)
    cu_out, _, tphys, tot = s.download()
    if not np.isfinite(cu_out).all():
        print("non-finite state after the timed steps", file=sys.stderr)
        sys.exit(1)
    active = int((m.iblank == 1).sum())
    updates = active * 4 * args.steps  # the timed steps only
    ms_per_step = total_ms / args.steps
    value = updates / (total_ms * 1e-3)
    flux_avg_ms = flux_ms / max(nflux, 1)
    bpc = FLUX_BYTES_RULE if args.precision == 0 else FLUX_BYTES_RULE_F32
    layout_bpc = FLUX_BYTES_PER_CELL if args.precision == 0 else FLUX_BYTES_PER_CELL_F32
    achieved = bpc * active / (flux_avg_ms * 1e-3) / 1e9
    out = {
        "metric": "cell-stage updates/s @5M tets, %HBM peak; serve p50/p99 solve latency",
        "value": value,
        "unit": "cell-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == 0 else "f32-bulk/f64-state",
        "data": "synthetic",
        "cell_steps_per_s": value / 4.0,  # whole 4-stage steps, comparable across RK variants
        "config": {"workload": (f"box {args.n[0]}x{args.n[1]}x{args.n[2]} hexes x6 tets cut at r=0.15 + "
                                "annulus 10x112 hexes x6 around the r=0.1 wall, overset donor lists"
                                if donors is not None else
                                f"box {args.n[0]}x{args.n[1]}x{args.n[2]} hexes x6 tets, cylinder wall hole r=0.1")
                               + ", box cells in 8x8 y-z slabs, 8 water spheres, viscous, "
                               + ("FP64" if args.precision == 0 else "FP32 bulk"),
                   "tets": m.neles, "active_cells": active, "ghosts": m.nghosts,
                   "receptors": 0 if donors is None else int(donors[0].size), "stages_per_step": 4},
        "roofline": {"bound": "hbm", "kernel": "k_flux", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (flux_traffic_bytes() if list(args.n) == [96, 96, 90] and args.precision == 0
                                 and not args.single_mesh else None),
                     "flux_avg_ms": flux_avg_ms, "bytes_per_cell": bpc,
                     "layout_bytes_per_cell": layout_bpc,
                     "frac_on_layout_bytes": layout_bpc * active / (flux_avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
    }
    s.close()
    # measured copy ceiling beside the published peak
    out["roofline"]["copy_ceiling"] = stream_copy_gbs()
    log = lambda msg: print(msg, file=sys.stderr, flush=True)
    if not args.no_ensemble:
        out["ensemble"] = ensemble_bench(log)
        log(f"# ensemble: {out['ensemble']['updates_per_s']:.3e} updates/s")
    if not args.no_serve:
        out["serve"] = serve_bench(log)
        log(f"# serve: p50 {out['serve']['tts_p50_s']:.2f} s p99 {out['serve']['tts_p99_s']:.2f} s")
        # the same request mix arriving below the GPU's capacity (about 2.8
        # requests/s for this mix): time to solution without the queueing
        # that 20 requests/s imposes
        out["serve_light"] = serve_bench(log, n_requests=16, rate=1.5, sample_frac=0.0, seed=8)
        log(f"# serve at 1.5/s: p50 {out['serve_light']['tts_p50_s']:.2f} s "
            f"p99 {out['serve_light']['tts_p99_s']:.2f} s")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(m, p, cu, gp, donors, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
