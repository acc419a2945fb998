"""The benchmark workload itself, 100 steps, against the oracle.

The 5,230,160-tet two-mesh case (96x96x90 hex box x6 jittered tets cut at
r = 0.15, cells in 8x8 y-z slabs, plus the 10x112x90-hex annulus around the
r = 0.1 wall; 224,877 receptors on fixed donor lists; 8 water spheres,
viscous gas, FP64, HLLC-Batten) is stepped on the device by hp_step (the
captured step graph, donor gather inside), as the benchmark steps it.  Its
conserved field and ghost primitives after 10 and after 100 steps must equal
the oracle's bit for bit: the SHA-256 digests in
tests/golden/overset96_100steps.json, written by
tests/golden/make_overset96_hash.py from a 100-step oracle run (about 50
minutes on one core).  On a mismatch the seeded sample cells stored there
show how far apart the two are.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "overset96_100steps.json")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_bench_case_100_steps_equal_oracle_hashes():
    from hydraphase import Solver
    want = json.load(open(GOLDEN))
    m, p, cu, gp, donors = bench.build_case(96, 96, 90, overset=True)
    assert m.neles == want["neles"] and donors[0].size == want["receptors"]
    idx = np.asarray(want["sample_idx"])
    print(f"bench case built: {m.neles} tets, {donors[0].size} receptors", flush=True)
    s = Solver(m, p)
    try:
        s.upload(cu, gp, 0.0, 0)
        s.set_donors(*donors)
        s.apply_donors()
        done = 0
        for target in (10, 100):
            s.step(target - done)
            done = target
            cu_d, gp_d, t, tot = s.download()
            rec = want[f"step{target}"]
            print(f"step {target}: t = {t!r}", flush=True)
            assert int((~np.isfinite(cu_d)).sum()) == rec["nonfinite"]
            if sha(cu_d) != rec["cu_sha256"]:
                ref = np.array([[float.fromhex(v) for v in row] for row in rec["sample"]])
                got = cu_d[:, idx]
                rel = np.abs(got - ref).max(axis=1) / np.abs(ref).max(axis=1)
                pytest.fail(f"step {target}: conserved field differs from the oracle's; "
                            f"sample max relative difference per variable {rel.tolist()}")
            assert sha(gp_d) == rec["gp_sha256"], f"step {target}: ghost primitives differ"
            assert (t, tot) == (rec["time"], rec["total"])
    finally:
        s.close()
