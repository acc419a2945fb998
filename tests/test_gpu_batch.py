"""Batches (hp_batch): many cases advanced by the same launches.

Every cell's arithmetic reads only its own case, so a case's result must be
bit-for-bit the same alone (hp_step) and in any mix of other cases, with or
without the captured graph, for any launch shape the tuner may pick, and
while other cases join and leave.  Each case runs its own dt.
"""
import numpy as np
import pytest

import oracle as O
from hydraphase import Solver
from hydraphase import batch as HB
from hydraphase import mesh as M

pytestmark = pytest.mark.gpu


def specs():
    # uneven sizes, both viscosities, different Mach numbers and radii
    return [HB.CaseSpec(5, 0.3, 0.2, True, 11), HB.CaseSpec(8, 0.6, 0.12, False, 12),
            HB.CaseSpec(4, 0.15, 0.25, True, 13), HB.CaseSpec(7, 0.45, 0.08, False, 14)]


def alone(case, nsteps):
    m, p, cu, gp, _ = case
    s = Solver(m, p)
    s.upload(cu, gp, 0.0, 0)
    s.step(nsteps)
    out = s.download()
    s.close()
    return out


def in_batch(cases, nsteps, graph=True, grid=(0, 0, 0), block=1):
    b = HB.Batch(*HB.capacity([c[0] for c in cases]))
    b.set_grid(*grid)
    b.set_block(block)
    sols = []
    for (m, p, cu, gp, _) in cases:
        s = Solver(m, p)
        s.upload(cu, gp, 0.0, 0)
        b.add(s)
        sols.append(s)
    b.step(nsteps, graph=graph)
    b.synchronize()
    outs = [s.download() for s in sols]
    b.close()
    for s in sols:
        s.close()
    return outs


def same(a, b):
    cu_a, gp_a, t_a, n_a = a
    cu_b, gp_b, t_b, n_b = b
    return n_a == n_b and t_a == t_b and np.array_equal(cu_a, cu_b) and np.array_equal(gp_a, gp_b)


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_equals_alone(precision):
    cases = [HB.build_case(s, precision) for s in specs()]
    ref = [alone(c, 4) for c in cases]
    for graph in (True, False):
        got = in_batch(cases, 4, graph=graph)
        for k in range(len(cases)):
            assert same(got[k], ref[k]), (precision, graph, k)


@pytest.mark.parametrize("myflux,myorder", [(1, 2), (2, 1), (1, 1)])
def test_batch_equals_alone_other_schemes(myflux, myorder):
    """HLLC-Zien and first-order faces in a batch: each case equals its
    alone run, FP64 and FP32 bulk."""
    for precision in (0, 1):
        cases = []
        for s in specs()[:3]:
            m, p, cu, gp, perm = HB.build_case(s, precision, myflux=myflux)
            p["myorder"] = myorder
            cases.append((m, p, cu, gp, perm))
        ref = [alone(c, 3) for c in cases]
        got = in_batch(cases, 3, graph=True)
        for k in range(len(cases)):
            assert same(got[k], ref[k]), (precision, k)


def test_batch_fp64_case_matches_oracle():
    cases = [HB.build_case(s, 0) for s in specs()[:2]]
    got = in_batch(cases, 3)
    for (m, p, cu, gp, _), g in zip(cases, got):
        o = O.Oracle(m, p)
        o.set_state(cu, gp, 0.0, 0)
        o.run(3)
        cu_o, gp_o, t_o, n_o = o.get_state()
        assert n_o == g[3] == 3 and t_o == g[2]
        assert np.array_equal(cu_o, g[0])


@pytest.mark.parametrize("precision", [0, 1])
def test_launch_shapes_give_same_results(precision):
    """Persistent grids (blocks looping over several tiles, with the flux's
    staged primitive recovery between tiles) equal one block per tile."""
    cases = [HB.build_case(s, precision) for s in specs()]
    ref = in_batch(cases, 3)
    for grid in [(8, 8, 8), (16, 24, 8), (512, 1024, 64)]:
        got = in_batch(cases, 3, grid=grid)
        for k in range(len(cases)):
            assert same(got[k], ref[k]), (grid, k)


@pytest.mark.parametrize("precision", [0, 1])
def test_block_sizes_give_same_results(precision):
    """Blocks of 2 or 4 tiles for the streaming per-cell kernels (one tile
    per wave), with one block per tile of the other families or persistent
    grids, eager and graph: the same bits as blocks of one tile."""
    cases = [HB.build_case(s, precision) for s in specs()]
    ref = in_batch(cases, 3)
    for block in (2, 4):
        for grid, graph in [((0, 0, 0), True), ((0, 0, 0), False), ((8, 8, 8), True), ((24, 16, 8), False)]:
            got = in_batch(cases, 3, graph=graph, grid=grid, block=block)
            for k in range(len(cases)):
                assert same(got[k], ref[k]), (block, grid, graph, k)
    b = HB.Batch(1, 1000, 1000)
    with pytest.raises(Exception):
        b.set_block(3)
    b.close()


def test_join_and_leave_with_one_graph():
    """A: steps 0-5; B joins at step 2 of A and leaves after 2 steps; C
    joins with B and stays.  Each must equal its alone run."""
    sp = specs()
    ca, cb, cc = (HB.build_case(s, 0) for s in sp[:3])
    b = HB.Batch(*HB.capacity([ca[0], cb[0], cc[0]]))
    sols = {}
    for name, c in (("a", ca), ("b", cb), ("c", cc)):
        s = Solver(c[0], c[1])
        s.upload(c[2], c[3], 0.0, 0)
        sols[name] = s
    b.add(sols["a"])
    b.step(2)
    b.add(sols["b"])
    b.add(sols["c"])
    b.step(2)
    b.remove(sols["b"])
    out_b = sols["b"].download()
    b.step(1)
    b.synchronize()
    out_a = sols["a"].download()
    out_c = sols["c"].download()
    b.close()
    assert same(out_a, alone(ca, 5))
    assert same(out_b, alone(cb, 2))
    assert same(out_c, alone(cc, 3))


def test_morton_order_same_per_cell_results():
    sp = specs()[1]
    nat = HB.build_case(sp, 0, morton=False)
    mor = HB.build_case(sp, 0, morton=True)
    a = alone(nat, 3)
    b = in_batch([mor], 3)[0]
    perm = mor[4]
    assert np.array_equal(b[0], a[0][:, perm])
    assert np.array_equal(b[1], a[1]) and b[2] == a[2]


def test_case_bytes_covers_allocation():
    """hp_case_bytes is what the request loop admits by: the device memory a
    case really takes must not exceed it by more than allocation rounding."""
    import ctypes
    from hydraphase import _lib
    _lib.lib()
    hip = ctypes.CDLL("libamdhip64.so.7")  # the HIP runtime the library already loaded

    def free_bytes():
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        assert hip.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
        return f.value

    m, p, cu, gp, _ = HB.build_case(HB.CaseSpec(32, 0.3, 0.2, True, 21), 1)
    est = HB.case_bytes(m, 1)
    free0 = free_bytes()
    s = Solver(m, p)
    s.synchronize()
    free1 = free_bytes()
    used = free0 - free1
    s.close()
    print(f"case bytes: estimate {est / 2**20:.1f} MiB, allocated {used / 2**20:.1f} MiB")
    assert used <= 1.05 * est + (64 << 20), (used, est)


def test_run_ensemble_matches_alone():
    """hydraphase.batch.run_ensemble (the ensemble driver: Morton-ordered
    cases, graph replay, each case leaving when its own step count is done)
    returns per case the state of that case run alone, and its health."""
    sp = [HB.CaseSpec(5, 0.3, 0.2, True, 21, steps=4), HB.CaseSpec(7, 0.5, 0.1, False, 22, steps=7),
          HB.CaseSpec(4, 0.2, 0.15, True, 23, steps=2)]
    res = HB.run_ensemble(sp, precision=0, tuning={"morton": True, "grid": (0, 0, 0)}, graph=True)
    assert res["health"] == [0, 0, 0]
    for k, s in enumerate(sp):
        m, p, cu, gp, _ = HB.build_case(s, 0)
        ref = alone((m, p, cu, gp, None), s.steps)
        got = res["states"][k]
        perm = res["perms"][k]
        cu_got = np.empty_like(got[0])
        cu_got[:, perm] = got[0]  # back to the generated cell order
        assert got[3] == ref[3] == s.steps and got[2] == ref[2]
        assert np.array_equal(cu_got, ref[0]) and np.array_equal(got[1], ref[1]), k
