"""Overset donor lists (hydraphase.overset) and the oracle's gather.

The reference has no donor interpolation, so these results are pinned by
what the construction fixes, not by reference vectors (parity unpinned):
partition of unity, exact reproduction of linear fields, one-hot weights
when the receptor sits on a donor centroid, and the gather formula itself
(a dropped or misplaced term fails the linear-field test)."""
import os

import numpy as np
import pytest

import oracle as O
from hydraphase import mesh as M
from hydraphase import overset as OV


@pytest.fixture(scope="module")
def small():
    # background 16^3 (h = 1/16), cut at r = 0.15, one fringe layer;
    # annulus 0.1..0.35 in 5 layers: the fringes lie inside each other's
    # active regions
    return OV.overset_mesh(16, 16, 16, nr=5, ntheta=24, r_outer=0.35, fringe=1)


def test_annulus_geometry():
    x, y, z, nod = M.annulus_mesh(4, 32, 6)
    p = [np.stack([x[nod[:, k]], y[nod[:, k]], z[nod[:, k]]], 1) for k in range(4)]
    v = np.einsum("ij,ij->i", np.cross(p[1] - p[0], p[2] - p[0]), p[3] - p[0]) / 6.0
    assert (v > 0).all()
    # a 32-gon shell: area pi (r1^2 - r0^2) * sin(2pi/32)/(2pi/32), length 1
    exact = np.pi * (0.25 ** 2 - 0.1 ** 2) * np.sin(2 * np.pi / 32) / (2 * np.pi / 32)
    assert abs(v.sum() - exact) < 1e-12
    m = M.combine([(x, y, z, nod)])
    assert set(np.unique(m.gtype)) == {M.BC_WALL, M.BC_Z}
    assert (m.nc < m.ntot).all() and (m.nc >= 0).all()


def test_donor_weights_partition_of_unity_and_linear(small):
    m, recv, donors, w = small
    c = OV.centroids(m)
    assert recv.size > 0 and np.all(m.iblank[recv] == 0) and np.all(m.iblank[donors] == 1)
    assert np.abs(w.sum(axis=1) - 1.0).max() < 1e-13
    # no extrapolation: every weight non-negative (to roundoff)
    assert (w >= -1e-12).all()
    rng = np.random.default_rng(4)
    for _ in range(3):
        a, b = rng.normal(), rng.normal(size=3)
        f = a + c @ b
        interp = (w * f[donors]).sum(axis=1)
        exact = np.abs(interp - f[recv]) < 1e-12 * (abs(a) + np.abs(b).sum())
        # exact for linear fields except beyond the hull of the other mesh's
        # donor centroids: the first and last cell layers in z
        assert exact.mean() > 0.9
        assert (np.abs(c[recv[~exact], 2]) > 0.5 - 1.0 / 16).all()
    # receptors of one mesh take donors from the other (annulus cells,
    # 5 x 24 x 16 hexes x 6, come after the background cells)
    nbg = m.neles - 5 * 24 * 16 * 6
    bg = recv < nbg
    assert bg.any() and (~bg).any()
    assert np.all(donors[bg] >= nbg) and np.all(donors[~bg] < nbg)


def test_donor_on_a_centroid_is_one_hot():
    rng = np.random.default_rng(1)
    pts = rng.uniform(-1, 1, size=(200, 3))
    ids = np.arange(200)
    d, w, _ = OV.find_donors(pts[:20], pts, ids)
    hot = np.abs(w - 1.0) < 1e-12
    assert hot.sum(axis=1).tolist() == [1] * 20
    assert np.all(d[hot] == ids[:20])
    assert np.abs(w[~hot]).max() < 1e-12


def test_donor_file_roundtrip(tmp_path, small):
    m, recv, donors, w = small
    p = os.path.join(tmp_path, "donor_proc0000.in")
    OV.write_donors(p, recv, donors, w)
    r2, d2, w2 = OV.read_donors(p)
    assert np.array_equal(r2, recv) and np.array_equal(d2, donors) and np.array_equal(w2, w)


def test_oracle_gather_reproduces_linear_conserved_fields(small):
    m, recv, donors, w = small
    p = M.water_air_params(cfl=0.5)
    cu, gp = M.sphere_state(m, p, [(0.3, 0.0, 0.0, 0.1)])
    o = O.Oracle(m, p)
    o.set_state(cu, gp, 0.0, 0)
    c = OV.centroids(m)
    cn = o.array("cn")
    rng = np.random.default_rng(2)
    lin = []
    for k in range(10):
        a, b = rng.normal(), rng.normal(size=3)
        cn[k, : m.neles] = a + c @ b
        lin.append((a, b))
    before = cn[:, : m.neles].copy()
    o.set_donors(recv, donors, w)
    o.call("donor_gather")
    # receptors whose donor weights reproduce their centroid (all but those
    # beyond the hull of the other mesh's donor centroids, whose weights
    # were clipped to a convex combination)
    inside = np.abs((w[:, :, None] * c[donors]).sum(axis=1) - c[recv]).max(axis=1) < 1e-12
    assert inside.mean() > 0.9
    for k, (a, b) in enumerate(lin):
        err = np.abs(cn[k, recv] - (a + c[recv] @ b))
        assert err[inside].max() < 1e-12 * (abs(a) + np.abs(b).sum())
        f = before[k]
        want = w[:, 0] * f[donors[:, 0]] + w[:, 1] * f[donors[:, 1]] + w[:, 2] * f[donors[:, 2]] + \
            w[:, 3] * f[donors[:, 3]]
        assert np.array_equal(cn[k, recv], want)


def test_slab_numbering_keeps_the_case():
    """Numbering the background in y-z slabs gives the same cells (same
    centroids), the same receptors with the same donors and weights up to
    the renumbering, the annulus in place, and brings y/z face neighbours
    closer."""
    m0, r0, d0, w0 = OV.overset_mesh(16, 16, 16, nr=5, ntheta=24, r_outer=0.35, fringe=1)
    m1, r1, d1, w1 = OV.overset_mesh(16, 16, 16, nr=5, ntheta=24, r_outer=0.35, fringe=1, slabs=(4, 4))
    c0, c1 = OV.centroids(m0), OV.centroids(m1)
    k0, k1 = np.lexsort(c0.T[::-1]), np.lexsort(c1.T[::-1])
    assert np.array_equal(c0[k0], c1[k1])
    new_of_old = np.empty(m0.neles, dtype=np.int64)
    new_of_old[k0] = k1
    pos1 = {int(r): k for k, r in enumerate(r1)}
    assert sorted(pos1) == sorted(int(x) for x in new_of_old[r0])
    for k in range(r0.size):
        j = pos1[int(new_of_old[r0[k]])]
        assert np.array_equal(new_of_old[d0[k]], d1[j]) and np.array_equal(w0[k], w1[j])
    nbg = m0.neles - 5 * 24 * 16 * 6  # the annulus (nr x ntheta x nz hexes x 6) comes last
    assert np.array_equal(new_of_old[nbg:], np.arange(nbg, m0.neles))

    def spread(m):
        nc = np.asarray(m.nc)
        i = np.arange(m.neles)
        d = np.concatenate([np.abs(nc[f] - i)[nc[f] < m.neles] for f in range(4)])
        return np.percentile(d, 90)
    assert spread(m1) < spread(m0)


def test_bench_case_donors_do_not_extrapolate():
    """The two-mesh bench workload (96x96x90 box cut at r = 0.15, annulus
    10 x 112 x 90 around the r = 0.1 wall, two fringe layers): every one of
    its 224,877 receptors has weights >= -1e-12 summing to one, so no
    receptor value leaves the range of its donors'; all but the few whose
    centroid lies beyond the hull of the other mesh's donor centroids (the
    first and last cell layers in z) interpolate linear fields exactly."""
    m, recv, donors, w = OV.overset_mesh(96, 96, 90, nr=10, ntheta=112, r_body=0.1, r_outer=0.25,
                                         r_cut=0.15, fringe=2, jitter=0.1, seed=1, slabs=(8, 8))
    assert recv.size == 224877
    assert (w >= -1e-12).all(), w.min()
    np.testing.assert_allclose(w.sum(axis=1), 1.0, rtol=0, atol=1e-14)
    c = OV.centroids(m)
    g = np.array([0.7, -1.3, 2.1])
    f = c @ g
    err = np.abs((w * f[donors]).sum(axis=1) - f[recv])
    exact = err <= 1e-12
    assert exact.mean() > 0.995, exact.mean()
    inexact = c[recv[~exact]]
    assert (np.abs(inexact[:, 2]) > 0.5 - 1.0 / 90).all()  # the end layers in z
