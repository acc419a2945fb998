"""Mesh and initial-state generator: the tables the reference's reader
expects (sr_geometry.F90:101-139) and the forward EOS relations."""
import numpy as np

from hydraphase import mesh as M
import oracle as O


def test_box_tables():
    m = M.box_mesh(5, 4, 3, jitter=0.1, seed=2, hole=(0.05, 0.0, 0.16))
    ne = m.neles
    # face neighbours are mutual
    for k in range(4):
        i = np.arange(ne)
        j = m.nc[k]
        inner = j < ne
        back = (m.nc[:, j[inner]] == i[inner]).any(axis=0)
        assert back.all()
    # ghosts: ids ne..ne+ng-1, each the neighbour of its parent's side
    g = np.arange(m.nghosts)
    assert (m.nc[m.gside - 1, m.gparent] == ne + g).all()
    assert set(np.unique(m.gtype)) == {2, 3, 5, 51, 53}
    # positive orientation (face k opposite node k, outward area vectors)
    p = [np.stack([m.x[m.nod[k]], m.y[m.nod[k]], m.z[m.nod[k]]]) for k in range(4)]
    v = np.einsum("ij,ij->j", p[1] - p[0], np.cross(p[2] - p[0], p[3] - p[0], axis=0))
    assert (v > 0).all()


def test_counts_kuhn_box():
    m = M.box_mesh(4, 3, 2, jitter=0.0)
    assert m.neles == 4 * 3 * 2 * 6
    assert m.nghosts == 2 * 2 * (4 * 3 + 4 * 2 + 3 * 2)
    types, counts = np.unique(m.gtype, return_counts=True)
    assert dict(zip(types, counts)) == {2: 2 * 3 * 2, 3: 2 * 3 * 2, 51: 2 * 2 * 4 * 2, 53: 2 * 2 * 4 * 3}


def test_permutation_roundtrip():
    m = M.box_mesh(4, 4, 4, jitter=0.1)
    perm = M.morton_order(m)
    mp = m.permuted(perm)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(m.neles)
    assert (mp.nod == m.nod[:, perm]).all()
    j = m.nc[:, perm]
    inner = j < m.neles
    assert (mp.nc[inner] == inv[j[inner]]).all()
    assert (mp.nc[~inner] == j[~inner]).all()
    assert (perm[mp.gparent] == m.gparent).all()


def test_sphere_state_inverts_through_eos():
    """The oracle's EOS inversion returns the primitives the generator
    used (pressure, temperature, velocity, phase fraction)."""
    m = M.box_mesh(6, 6, 6, jitter=0.1)
    p = M.water_air_params()
    cu, gp = M.sphere_state(m, p, [(0.0, 0.0, 0.0, 0.2)])
    o = O.Oracle(m, p)
    o.set_state(cu, gp, 0.0, 0)
    o.call("primitive")
    ne = m.neles
    np.testing.assert_allclose(o.q("pg")[:ne], p["pmin"], rtol=1e-9)
    np.testing.assert_allclose(o.q("tg")[:ne], p["tgmin"], rtol=1e-9)
    np.testing.assert_allclose(o.q("ug")[:ne], p["ugmin"], rtol=1e-9)
    phig = o.q("phig")[:ne]
    assert phig.min() < 0.5 and phig.max() > 0.99
