"""Pin the oracle's 100-step field of the 5.2M-tet two-mesh workload.

The 96x96x90 hex box (x6 jittered tets, seed 1, cut at r = 0.15, cells in
8x8 y-z slabs) plus the 10x112x90-hex near-body annulus around the r = 0.1
wall, coupled by fixed donor lists (two fringe layers): 5,230,160 tets,
224,877 receptors; 8 water spheres in air at Mach 0.3, viscous gas, CFL
0.5, FP64, myorder 2, HLLC-Batten.  The oracle runs 100 steps (about 35
minutes on one core) and the script writes tests/golden/overset96_100steps.json:
SHA-256 of the conserved field (10 x neles float64, C order) and of the
ghost primitives, the time and per-variable norms at steps 10 and 100, and
the values of 4,096 seeded sample cells (all ten variables) at both steps,
for a relative comparison should the hashes differ.

usage: python tests/golden/make_overset96_hash.py
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "multiphase-code_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)
import bench  # noqa: E402
import oracle as O  # noqa: E402

NSAMPLE = 4096


def record(cu, gp, t, tot, idx):
    fin = np.isfinite(cu)
    return {"time": t, "total": tot,
            "cu_sha256": hashlib.sha256(np.ascontiguousarray(cu).tobytes()).hexdigest(),
            "gp_sha256": hashlib.sha256(np.ascontiguousarray(gp).tobytes()).hexdigest(),
            "nonfinite": int((~fin).sum()),
            "cu_l2": [float(v) for v in np.sqrt((np.where(fin, cu, 0.0) ** 2).sum(axis=1))],
            "sample": [[float(v).hex() for v in row] for row in cu[:, idx]]}


def main():
    m, p, cu0, gp0, (recv, ids, w) = bench.build_case(96, 96, 90, overset=True)
    idx = np.sort(np.random.default_rng(7).choice(m.neles, NSAMPLE, replace=False))
    o = O.Oracle(m, p)
    o.set_state(cu0, gp0, 0.0, 0)
    o.set_donors(recv, ids, w)
    o.call("donor_gather")
    out = {"neles": m.neles, "receptors": int(recv.size), "sample_idx": idx.tolist()}
    t0 = time.time()
    done = 0
    for target in (10, 100):
        for _ in range(target - done):
            o.step()
            done += 1
            print(f"step {done} {time.time() - t0:.0f} s", flush=True)
        cu, gp, t, tot = o.get_state()
        out[f"step{target}"] = record(cu, gp, t, tot, idx)
    out["oracle_seconds"] = time.time() - t0
    path = os.path.join(ROOT, "tests", "golden", "overset96_100steps.json")
    json.dump(out, open(path, "w"), indent=0)
    print("wrote", path)


if __name__ == "__main__":
    main()
