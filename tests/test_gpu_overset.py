"""Overset donor gather on the GPU (hp_set_donors / hp_apply_donors) and
full steps of a two-mesh case against the oracle with the same lists."""
import numpy as np
import pytest

import oracle as O
from hydraphase import Solver
from hydraphase import mesh as M
from hydraphase import overset as OV

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    m, recv, donors, w = OV.overset_mesh(16, 16, 16, nr=5, ntheta=24, r_outer=0.35, fringe=1)
    p = M.water_air_params(cfl=0.5, viscous_gas=True)
    cu, gp = M.sphere_state(m, p, [(0.3, 0.05, 0.0, 0.12)])
    return m, p, cu, gp, recv, donors, w


def test_gather_matches_formula(case):
    m, p, cu, gp, recv, donors, w = case
    s = Solver(m, p)
    rng = np.random.default_rng(5)
    cn = rng.uniform(0.5, 2.0, size=(10, m.neles))
    s.set("cn", cn)
    s.set_donors(recv, donors, w)
    s.apply_donors()
    got = s.get("cn")
    want = cn.copy()
    for k in range(10):
        c = cn[k]
        want[k, recv] = w[:, 0] * c[donors[:, 0]] + w[:, 1] * c[donors[:, 1]] + w[:, 2] * c[donors[:, 2]] + \
            w[:, 3] * c[donors[:, 3]]
    assert np.array_equal(got, want)
    s.close()


@pytest.mark.parametrize("flux", [2, 1])
def test_overset_steps_match_oracle(case, flux):
    m, p, cu, gp, recv, donors, w = case
    p = dict(p, myflux=flux)
    s = Solver(m, p)
    s.upload(cu, gp, 0.0, 0)
    s.set_donors(recv, donors, w)
    s.apply_donors()
    s.step(1)                   # the captured step
    d1 = s.last_diffs()
    d23 = s.step_diffs(2)       # eager steps returning their monitor
    cu_g, gp_g, t_g, n_g = s.download()
    s.close()
    o = O.Oracle(m, p)
    o.set_state(cu, gp, 0.0, 0)
    o.set_donors(recv, donors, w)
    o.call("donor_gather")
    want = [o.step() for _ in range(3)]
    cu_o, gp_o, t_o, n_o = o.get_state()
    # the monitor (main.F90:603-612) over every cell, receptors included
    assert np.array_equal(d1, want[0]) and np.array_equal(d23, np.array(want[1:]))
    assert n_g == n_o == 3 and t_g == t_o
    assert np.isfinite(cu_g).all()
    assert np.array_equal(cu_g, cu_o)
    assert np.array_equal(gp_g, gp_o)


def test_donor_lists_rejected_when_bad(case):
    from hydraphase._lib import HydraError
    m, p, cu, gp, recv, donors, w = case
    s = Solver(m, p)
    bad = donors.copy()
    bad[0, 0] = recv[1]  # a receptor used as donor
    with pytest.raises(HydraError):
        s.set_donors(recv, bad, w)
    s.close()


def test_cleared_donor_lists_step_as_none(case):
    """hp_set_donors with no receptors drops the lists (and the fused step's
    donor flags and staging): the case then steps exactly as one that never
    had donor lists."""
    m, p, cu, gp, recv, donors, w = case
    s = Solver(m, p)
    s.upload(cu, gp, 0.0, 0)
    s.set_donors(recv, donors, w)
    s.set_donors(np.zeros(0, dtype=np.int32), np.zeros((0, 4), dtype=np.int32), np.zeros((0, 4)))
    s.step(2)
    got = s.download()
    s.close()
    s = Solver(m, p)
    s.upload(cu, gp, 0.0, 0)
    s.step(2)
    ref = s.download()
    s.close()
    assert got[3] == ref[3] == 2 and got[2] == ref[2]
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])


def run_overset(case, precision, steps, noise=0.0):
    from hydraphase import serve as SV
    m, p, cu, gp, recv, donors, w = case
    s = Solver(m, dict(p, precision=precision))
    s.upload(SV.perturbed(cu, noise) if noise else cu, gp, 0.0, 0)
    s.set_donors(recv, donors, w)
    s.apply_donors()
    s.step(steps)
    out = s.download()[0]
    s.close()
    return out


def test_fp32_bulk_overset_within_floor(case):
    """The FP32-bulk form on the two-mesh case (faces computed once, the
    donors' stage updates and the receptors' gather in FP64): after 10
    steps each conserved variable's relative L2 error against the FP64 path
    (bitwise the oracle's, above) stays within 4x that variable's
    conditioning floor, the FP64 path's own departure from a start perturbed
    by 2^-24 relative noise, and every value is finite."""
    ref = run_overset(case, 0, 10)
    f32 = run_overset(case, 1, 10)
    floor = run_overset(case, 0, 10, noise=2.0 ** -24)
    assert np.isfinite(f32).all()
    norm = np.sqrt((ref ** 2).sum(axis=1))
    err = np.sqrt(((f32 - ref) ** 2).sum(axis=1)) / norm
    flo = np.sqrt(((floor - ref) ** 2).sum(axis=1)) / norm
    print("fp32 rel L2", np.array2string(err, precision=2))
    print("floor rel L2", np.array2string(flo, precision=2))
    assert (err <= 4.0 * flo).all(), (err, flo)
