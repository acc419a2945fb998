"""Pin the oracle's states of the largest cases of the batched-ensemble and
serving workloads by hash, for the GPU tests that run those workloads at
their sizes.

* Ensemble (64 boxes of n^3 x 6 jittered tets, n log-uniform 8..48, Mach
  0.1-0.8, water sphere radius 0.05-0.25, half of them viscous; seed 1):
  its two largest cases, 20 steps each, FP64, HLLC-Batten, cells in the
  generated order.
* Serving mix (boxes n 10..70): one n = 70 box (2,058,000 tets), the
  serving test's largest request, 20 steps.

Writes tests/golden/workload_hashes.json: per case the spec, SHA-256 of the
conserved field (10 x neles float64, C order) and of the ghost primitives,
the time, and whether the state is finite.

usage: python tests/golden/make_workload_hashes.py
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
from hydraphase import batch as HB  # noqa: E402
import oracle as O  # noqa: E402

STEPS = 20


def serve_big_spec():
    """The serving test's n = 70 request (spec fields as the ensemble's)."""
    return HB.CaseSpec(70, 0.35, 0.08, True, 7070, STEPS)


def pin(spec):
    m, p, cu, gp, _ = HB.build_case(spec, 0)
    o = O.Oracle(m, p)
    o.set_state(cu, gp, 0.0, 0)
    t0 = time.time()
    o.run(STEPS)
    cu1, gp1, t, tot = o.get_state()
    print(f"n={spec.n} ({m.neles} tets): {time.time() - t0:.0f} s", flush=True)
    return {"spec": [spec.n, spec.mach, spec.radius, spec.viscous, spec.seed], "neles": m.neles,
            "steps": STEPS, "time": t, "total": tot, "finite": bool(np.isfinite(cu1).all()),
            "cu_sha256": hashlib.sha256(np.ascontiguousarray(cu1).tobytes()).hexdigest(),
            "gp_sha256": hashlib.sha256(np.ascontiguousarray(gp1).tobytes()).hexdigest()}


def main():
    specs = HB.ensemble_specs(64, seed=1)
    order = np.argsort([-s.n for s in specs], kind="stable")[:2]
    out = {"ensemble": {str(int(k)): pin(specs[k]) for k in order}, "serve_n70": pin(serve_big_spec())}
    path = os.path.join(ROOT, "tests", "golden", "workload_hashes.json")
    json.dump(out, open(path, "w"), indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
