"""HIP path (libhydra_hip.so) against the CPU oracle, through the C ABI.

Every routine is compared from an identical input state; the FP64 path
evaluates the reference's expressions in the reference's order without
contraction, so the bar is bit-for-bit equality.  The full-step tests bound
the relative L2 difference of the conserved variables by 1e-10 (the
drop-in tolerance) and report whether it is exactly zero.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import oracle as O
from hydraphase import mesh as M

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_breakdown_steps_bitwise():
    """A case whose state breaks down (the 10 x 9 x 8 box with the wall
    cylinder: 34 cells hold NaN primitives after two steps, as the
    reference's own step gives them).  The CFL minimum over NaN bounds
    depends on the order it is taken in; the GPU folds its tiles in cell
    order as the reference's MIN loop does, so eight steps still equal the
    oracle's (NaN where the oracle has NaN), and so does hp_dt_min on the
    broken state."""
    m = M.box_mesh(10, 9, 8, jitter=0.1, seed=3, hole=(0.05, 0.0, 0.16))
    p = M.water_air_params(cfl=0.5, viscous_gas=True)
    cu0, gp0 = M.sphere_state(m, p, [(0.0, 0.1, 0.0, 0.22)])
    s = solver_cls()(m, p)
    o = O.Oracle(m, p)
    try:
        s.upload(cu0, gp0, 0.0, 0)
        o.set_state(cu0, gp0, 0.0, 0)
        broke = False
        for k in range(8):
            s.step(1)
            o.run(1)
            cu, gp, t, tot = s.download()
            cu_o, gp_o, t_o, tot_o = o.get_state()
            assert tot == tot_o and (t == t_o or (np.isnan(t) and np.isnan(t_o))), (k, t, t_o)
            assert_same(f"cu step {k + 1}", cu, cu_o)
            assert_same(f"ghost prims step {k + 1}", gp, gp_o)
            broke = broke or bool(np.isnan(cu).any())
        assert broke, "the case was meant to break down"
        s.upload(cu0, gp0, 0.0, 0)
        o.set_state(cu0, gp0, 0.0, 0)
        s.step(2)
        o.run(2)
        s.primitive()
        o.call("primitive")
        assert np.isnan(o.array("q")[:14, :m.neles]).any()
        d_g, d_o = s.dt_min(), o.call("dtmin")
        assert d_g == d_o or (np.isnan(d_g) and np.isnan(d_o)), (d_g, d_o)
    finally:
        s.close()


def solver_cls():
    from hydraphase import Solver
    return Solver


def small_case(flux=2, seed=5, steps=1, hole=True, blank=True):
    m = M.box_mesh(6, 5, 4, jitter=0.1, seed=seed, hole=(0.05, 0.0, 0.16) if hole else None)
    if blank:
        ib = m.iblank.copy()
        ib[[7, 50, 123]] = 0
        m.iblank = ib
    p = M.water_air_params(cfl=0.5, viscous_gas=True, myflux=flux)
    cu, gp = M.sphere_state(m, p, [(0.0, 0.1, 0.0, 0.22)])
    rng = np.random.default_rng(seed)
    for k in (1, 2, 3):
        cu[k] += cu[0] * rng.uniform(-30.0, 30.0, m.neles)
        cu[5 + k] += cu[5] * rng.uniform(-30.0, 30.0, m.neles)
    gp[1] = rng.uniform(-20.0, 20.0, m.nghosts)
    gp[2] = rng.uniform(-20.0, 20.0, m.nghosts)
    o = O.Oracle(m, p)
    o.set_state(cu, gp, 0.0, 0)
    for _ in range(steps):
        o.step()
    return m, p, o


def copy_state_to_gpu(o, s, m):
    ne = m.neles
    s.set("cn", o.array("cn")[:, :ne])
    s.set("cu", o.array("cu")[:, :ne])
    s.set("cn1", o.array("cn1")[:, :ne])
    s.set("q", o.array("q")[:14])
    g = o.array("grad")[:, :ne].reshape(13, 3, ne)
    s.set("grad", np.concatenate([g[:10], g[12:13]]).reshape(33, ne))
    s.set("phi", o.array("phi")[:ne])
    s.set("vis", o.array("vis"))
    s.set("d24", o.array("d24"))
    s.set("d2q", o.array("d2q")[:, :ne])
    s.set("anu", o.array("anu")[:, :ne])
    s.set("rhs", o.array("rhs"))
    _, _, t, tot = o.get_state()
    s.set_dt(o.dt, t, tot)


def assert_same(name, got, ref):
    bad = ~((got == ref) | (np.isnan(got) & np.isnan(ref)))
    if bad.any():
        idx = np.argwhere(bad)[:5]
        rel = np.abs(got - ref)[bad] / np.maximum(np.abs(ref[bad]), 1e-300)
        raise AssertionError(f"{name}: {bad.sum()} of {bad.size} differ; first {idx.tolist()}; "
                             f"max rel {rel.max():.3e}")


@pytest.fixture(scope="module")
def case():
    m, p, o = small_case(steps=1)
    o.call("primitive")
    _, _, t, tot = o.get_state()
    dt = o.call("dtmin") * p["cfl"]
    o.set_dt(dt, t + dt, tot + 1)
    o.call("smooth")
    o.call("viscousgp")
    o.call("derivative")
    o.call("venkat")
    return m, p, o


def fresh(case):
    m, p, o = case
    s = solver_cls()(m, p)
    copy_state_to_gpu(o, s, m)
    return m, p, o, s


def snapshot(o):
    return {n: o.array(n).copy() for n in ["cn", "cu", "cn1", "d2q", "rhs", "d24", "anu", "q", "grad", "phi", "vis"]}


def restore(o, snap, dt_t_tot):
    for n, v in snap.items():
        o.array(n)[...] = v
    o.set_dt(*dt_t_tot)


@pytest.mark.parametrize("routine", ["primitive", "dt", "smooth", "viscousgp", "flux", "rk_stage"])
def test_routine_bitwise(case, routine):
    m, p, o = case
    snap = snapshot(o)
    _, _, t, tot = o.get_state()
    dtt = (o.dt, t, tot)
    ne = m.neles
    try:
        m, p, o, s = fresh(case)
        if routine == "primitive":
            o.call("primitive"); s.primitive()
            assert_same("cn", s.get("cn"), o.array("cn")[:, :ne])
            assert_same("q", s.get("q")[:, :ne], o.array("q")[:14, :ne])
        elif routine == "dt":
            assert s.dt_min() == o.call("dtmin")
        elif routine == "smooth":
            o.call("smooth"); s.smooth()
            assert_same("anu", s.get("anu"), o.array("anu")[:, :ne])
            assert_same("d2q", s.get("d2q"), o.array("d2q")[:, :ne])
            assert_same("d24", s.get("d24"), o.array("d24"))
        elif routine == "viscousgp":
            o.call("viscousgp"); s.viscousgp()
            live = m.iblank == 1
            assert_same("vis", s.get("vis")[:, live], o.array("vis")[:, live])
            assert_same("ghost q", s.get("q")[:, ne:], o.array("q")[:14, ne:])
        elif routine == "flux":
            for flux in (2, 1):
                restore(o, snap, dtt)
                if flux == 1:
                    p1 = dict(p, myflux=1)
                    o1 = O.Oracle(m, p1)
                    restore(o1, snap, dtt)
                    s1 = solver_cls()(m, p1)
                    copy_state_to_gpu(o1, s1, m)
                    oo, ss = o1, s1
                else:
                    oo, ss = o, s
                oo.call("flux_solver"); ss.flux()
                live = m.iblank == 1
                assert_same(f"rhs myflux={flux}", ss.get("rhs")[:, live], oo.array("rhs")[:, live])
                assert_same(f"cn myflux={flux}", ss.get("cn"), oo.array("cn")[:, :ne])
                assert_same(f"q myflux={flux}", ss.get("q")[:, :ne], oo.array("q")[:14, :ne])
                assert_same(f"ghost q myflux={flux}", ss.get("q")[:, ne:], oo.array("q")[:14, ne:])
                g = oo.array("grad")[:, :ne].reshape(13, 3, ne)
                gref = np.concatenate([g[:10], g[12:13]]).reshape(33, ne)
                assert_same(f"grad myflux={flux}", ss.get("grad")[:, live], gref[:, live])
                assert_same(f"phi myflux={flux}", ss.get("phi")[live], oo.array("phi")[:ne][live])
        elif routine == "rk_stage":
            o.call("flux_solver"); s.flux()
            for st in (1, 2, 3, 4):
                o.call("rk_stage", st); s.rk_stage(st)
                assert_same(f"cn stage {st}", s.get("cn"), o.array("cn")[:, :ne])
            assert_same("cn1", s.get("cn1"), o.array("cn1")[:, :ne])
    finally:
        restore(o, snap, dtt)


def run_pair(m, p, o, s, nsteps, fused=True):
    """nsteps on both; returns GPU and oracle restart states."""
    if fused:
        s.step(nsteps)
    else:
        for _ in range(nsteps):
            s.primitive()
            dtmin = s.dt_min()
            _, _, t, tot = s.download()
            dt = dtmin * p["cfl"]
            s.set_dt(dt, t + dt, tot + 1)
            s.smooth()
            s.rk()
            s.end_step()
    o.run(nsteps)
    return s.download(), o.get_state()


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("flux", [2, 1])
def test_steps_bitwise(flux, fused):
    m, p, o = small_case(flux=flux, steps=0)
    s = solver_cls()(m, p)
    cu0, gp0, t0, tot0 = o.get_state()
    s.upload(cu0, gp0, t0, tot0)
    (cu, gp, t, tot), (cur, gpr, tr, totr) = run_pair(m, p, o, s, 5, fused=fused)
    assert tot == totr == 5
    assert t == tr
    assert_same("cu", cu, cur)
    assert_same("ghost prims", gp, gpr)


@pytest.mark.parametrize("flux", [2, 1])
def test_graph_replays_timing_and_monitor(flux):
    """hp_step replays one captured step: chunks of 1 and 2 replays with
    host reads in between, then eager steps with per-flux timing and with the
    per-step monitor, all equal the oracle step for step, and the device
    monitor of each step equals the oracle's max|cn-cu| (main.F90:603-612).
    Five steps: this case's reference step breaks down (NaN) at step 6 with
    HLLC-Zien, where a maximum over NaN depends on the order it is taken in
    (as Fortran's MAX with a NaN argument is processor dependent)."""
    m, p, o = small_case(flux=flux, steps=0)
    s = solver_cls()(m, p)
    cu0, gp0, t0, tot0 = o.get_state()
    s.upload(cu0, gp0, t0, tot0)
    done = 0
    for chunk in (1, 2):
        s.step(chunk)
        want = [o.step() for _ in range(chunk)]
        done += chunk
        got = s.download()
        ref = o.get_state()
        assert got[3] == ref[3] == done and got[2] == ref[2]
        assert_same(f"cu after {done}", got[0], ref[0])
        assert_same(f"ghosts after {done}", got[1], ref[1])
        assert np.array_equal(s.last_diffs(), want[-1]), (done, s.last_diffs(), want[-1])
    s.set_flux_timing(True)
    s.step(1)
    want = o.step()
    tot, fl, nfl = s.last_timing()
    assert nfl == 4 and 0.0 < fl <= tot
    assert np.array_equal(s.last_diffs(), want)
    s.set_flux_timing(False)
    d = s.step_diffs(1)
    want = np.array([o.step()])
    assert np.array_equal(d, want)
    got, ref = s.download(), o.get_state()
    assert np.isfinite(ref[0]).all()
    assert got[3] == ref[3] == done + 2 and got[2] == ref[2]
    assert_same("cu at the end", got[0], ref[0])
    s.close()


@pytest.mark.parametrize("flux", [2, 1])
def test_first_order_steps_bitwise(flux):
    """myorder = 1: no reconstruction; the interphase source still comes
    from its own kernel (no gradients or limiter are computed)."""
    m = M.box_mesh(6, 5, 4, jitter=0.1, seed=5, hole=(0.05, 0.0, 0.16))
    p = M.water_air_params(cfl=0.5, viscous_gas=True, myflux=flux)
    p["myorder"] = 1
    cu0, gp0 = M.sphere_state(m, p, [(0.0, 0.1, 0.0, 0.22)])
    o = O.Oracle(m, p)
    o.set_state(cu0, gp0, 0.0, 0)
    s = solver_cls()(m, p)
    s.upload(cu0, gp0, 0.0, 0)
    for fused in (True, False):
        (cu, gp, t, tot), (cur, gpr, tr, totr) = run_pair(m, p, o, s, 3, fused=fused)
        assert tot == totr and t == tr
        assert_same(f"cu fused={fused}", cu, cur)
        assert_same(f"ghost prims fused={fused}", gp, gpr)


def test_box24_100_steps():
    """The 24^3 x 6 jittered-box workload: 82,944 tets,
    water sphere R=0.2 in an air stream at Mach 0.3, CFL 0.5, 100 steps.
    The oracle's 100-step field is pinned by its SHA-256
    (tests/golden/box24_100steps.json, written by tests/golden/make_box24_hash.py);
    if the GPU field's hash differs, the oracle runs live and the relative
    L2 per conserved variable must stay <= 1e-10."""
    spec = json.load(open(os.path.join(HERE, "golden", "box24_100steps.json")))
    m = M.box_mesh(24, 24, 24, jitter=0.1, seed=1)
    p = M.water_air_params(cfl=0.5)
    cu0, gp0 = M.sphere_state(m, p, [(0.0, 0.0, 0.0, 0.2)])
    s = solver_cls()(m, p)
    s.upload(cu0, gp0, 0.0, 0)
    s.step(100)
    cu, gp, t, tot = s.download()
    assert tot == 100
    h = hashlib.sha256(np.ascontiguousarray(cu).tobytes()).hexdigest()
    print(f"box24 100 steps: gpu sha256 {h[:16]} oracle {spec['cu_sha256'][:16]} time {t!r}")
    if h == spec["cu_sha256"]:
        assert t == spec["time"]
        return
    o = O.Oracle(m, p)
    o.set_state(cu0, gp0, 0.0, 0)
    o.run(100)
    cur, _, tr, _ = o.get_state()
    rel = np.sqrt(((cu - cur) ** 2).sum(axis=1) / (cur ** 2).sum(axis=1))
    print("relative L2 per variable:", rel)
    assert np.all(rel <= 1e-10), rel


@pytest.mark.parametrize("flux", [2, 1])
def test_large_box_steps_bitwise(flux):
    """48^3 x 6 = 663,552 tets (10,368 tiles spread over every XCD's share
    of the grid, 27,648 boundary ghosts), water sphere R=0.2 in air: three
    fused steps equal the oracle's bit for bit, HLLC-Batten and
    HLLC-Zien."""
    m = M.box_mesh(48, 48, 48, jitter=0.1, seed=1)
    p = M.water_air_params(cfl=0.5, viscous_gas=True, myflux=flux)
    cu0, gp0 = M.sphere_state(m, p, [(0.0, 0.1, 0.0, 0.2)])
    s = solver_cls()(m, p)
    try:
        s.upload(cu0, gp0, 0.0, 0)
        s.step(3)
        cu, gp, t, tot = s.download()
    finally:
        s.close()
    o = O.Oracle(m, p)
    o.set_state(cu0, gp0, 0.0, 0)
    o.run(3)
    cu_o, gp_o, t_o, tot_o = o.get_state()
    assert tot == tot_o == 3 and t == t_o
    assert_same("cu", cu, cu_o)
    assert_same("ghost prims", gp, gp_o)


def test_bench_case_step_bitwise():
    """The benchmark's own workload at full size: the 96x96x90 box cut at
    r = 0.15 plus the near-body annulus, coupled by donor lists (5,230,160
    tets, 224,877 receptors), 8 water spheres, viscous, FP64.  One fused
    step (with the donor gather after every stage) equals the oracle's bit
    for bit."""
    import bench
    m, p, cu0, gp0, (recv, ids, w) = bench.build_case(96, 96, 90, overset=True)
    print(f"bench case: {m.neles} tets, {m.nghosts} ghosts, {recv.size} receptors", flush=True)
    s = solver_cls()(m, p)
    try:
        s.upload(cu0, gp0, 0.0, 0)
        s.set_donors(recv, ids, w)
        s.apply_donors()
        s.step(1)
        cu, gp, t, tot = s.download()
    finally:
        s.close()
    o = O.Oracle(m, p)
    o.set_state(cu0, gp0, 0.0, 0)
    o.set_donors(recv, ids, w)
    o.call("donor_gather")
    o.run(1)
    cur, gpr, tr, totr = o.get_state()
    assert tot == totr == 1 and t == tr
    assert np.isfinite(cu).all()
    assert_same("cu", cu, cur)
    assert_same("ghost prims", gp, gpr)


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 1, 1), (3, 2, 1)])
def test_meshes_below_one_tile_bitwise(dims):
    """Meshes of 6, 12 and 36 tets, one partial tile (most lanes of the
    block idle, every cell at a boundary): three steps, fused and per
    routine, equal the oracle's bit for bit."""
    m = M.box_mesh(*dims, jitter=0.1, seed=3)
    p = M.water_air_params(cfl=0.5, viscous_gas=True)
    cu0, gp0 = M.sphere_state(m, p, [(0.0, 0.0, 0.0, 0.3)])
    for fused in (True, False):
        o = O.Oracle(m, p)
        o.set_state(cu0, gp0, 0.0, 0)
        s = solver_cls()(m, p)
        try:
            s.upload(cu0, gp0, 0.0, 0)
            (cu, gp, t, tot), (cur, gpr, tr, totr) = run_pair(m, p, o, s, 3, fused=fused)
        finally:
            s.close()
        assert tot == totr == 3 and t == tr
        assert_same(f"cu fused={fused}", cu, cur)
        assert_same(f"ghost prims fused={fused}", gp, gpr)
