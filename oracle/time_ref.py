"""Time the reference's own compiled routines (oracle/_ref, amdflang -O2
-ffp-contract=off, unmodified sources) beside the C oracle (gcc -O2
-ffp-contract=off) on the same CPU core, routine by routine, on the 24^3 x 6
jittered box (82,944 tets) with a water sphere and a seeded momentum
perturbation after 3 steps.  Each routine runs REPS times from the same
input in one process (the harness times its own loop, excluding its file
I/O); whole first-order steps run the composed step of harness kernel 9
against the oracle's hyo_step.  Test infrastructure: it runs in the build
container, where the reference checkout is present, and prints a markdown
table of microseconds per call and the ratio.

usage: python oracle/time_ref.py [REPS]
"""
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "multiphase-code_amd"))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import make_golden as G  # noqa: E402
import oracle as O  # noqa: E402
from hydraphase import mesh as M  # noqa: E402

ORACLE_CALL = {"primitive": "primitive", "primitiveplusbc": "primitiveplusbc", "smooth": "smooth",
               "viscousgp": "viscousgp", "venkatakrishnan": "venkat", "source": "source",
               "hllcbatten": "flux_solver", "hllczien": "flux_solver"}


def main():
    reps = int(sys.argv[1]) if len(sys.argv) > 1 else 20
    if not os.path.exists(G.HARNESS):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)
    m = M.box_mesh(24, 24, 24, jitter=0.1, seed=1)
    rows = []
    for k, kid in G.KERNELS.items():
        myorder = 1 if k in ("hllcbatten", "hllczien", "steps") else 2
        flux = 1 if k == "hllczien" else 2
        p = M.water_air_params(cfl=0.5, viscous_gas=True, myflux=flux)
        p["myorder"] = myorder
        o = G.seeded_state(m, p, 5, 3)
        o.call("derivative")
        base = G.snapshot(o)
        n = 3 if k == "steps" else reps
        with tempfile.TemporaryDirectory() as td:
            fi, fo = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
            G.write_input(fi, o, base, p, kid, myorder, n)
            subprocess.run(["taskset", "-c", "0", G.HARNESS, fi, fo], check=True)
            t_ref = G.read_output(fo, m)["elapsed"] / n
        if k == "steps":
            cu, gp, t, tot = o.get_state()
            o2 = O.Oracle(m, p)
            o2.set_state(cu, gp, t, tot)
            t0 = time.perf_counter()
            for _ in range(n):
                o2.step()
            t_c = (time.perf_counter() - t0) / n
        else:
            t0 = time.perf_counter()
            for _ in range(n):
                o.call(ORACLE_CALL[k])
            t_c = (time.perf_counter() - t0) / n
        rows.append((k, t_ref * 1e6, t_c * 1e6))
        print(f"{k}: reference {t_ref * 1e3:.2f} ms, C oracle {t_c * 1e3:.2f} ms", file=sys.stderr, flush=True)
    print(f"| routine (82,944 tets) | reference (amdflang -O2) us/call | C oracle (gcc -O2) us/call | oracle / reference |")
    print("|---|---|---|---|")
    for k, a, b in rows:
        name = "first-order step (composed)" if k == "steps" else k
        print(f"| {name} | {a:.0f} | {b:.0f} | {b / a:.2f} |")


if __name__ == "__main__":
    main()
