"""The reference's file formats round-trip through hydraphase.refio."""
import os
import numpy as np

from hydraphase import mesh as M
from hydraphase import refio


def test_case_files_roundtrip(tmp_path):
    m = M.box_mesh(3, 3, 2, jitter=0.1, seed=4, hole=(0.0, 0.0, 0.12))
    p = M.water_air_params(viscous_gas=True)
    cu, gp = M.sphere_state(m, p, [(0.0, 0.0, 0.0, 0.2)])
    refio.write_case(str(tmp_path), m, p, cu, gp, time=1.25e-4)
    m2 = refio.read_mesh(str(tmp_path / "geometry_files"))
    for f in ("x", "y", "z", "nod", "nc", "iblank", "gparent", "gtype", "gside"):
        np.testing.assert_array_equal(getattr(m2, f), getattr(m, f), err_msg=f)
    t, cu2, gp2 = refio.read_restart(str(tmp_path / "restart_files" / "restart0000.in"), m.neles, m.nghosts)
    assert t == 1.25e-4
    np.testing.assert_array_equal(cu2, cu)
    np.testing.assert_array_equal(gp2, gp)
    q = refio.read_flow_in(str(tmp_path / "flow.in"))
    for k in ("pmax", "pinf", "gamap", "cpp", "epslnmin", "phipzero", "cfl", "diap", "as_", "ts"):
        assert q[k] == p[k], k
    assert q["istart"] == 1 and q["myflux"] == p["myflux"] and q["myorder"] == 2


def test_flow_in_reader_follows_reference_rules(tmp_path):
    """Blank and '#' lines are skipped, trailing comments dropped
    (sr_Read_flow.F90:19-32)."""
    p = M.water_air_params()
    refio.write_flow_in(str(tmp_path / "f.in"), p)
    txt = open(tmp_path / "f.in").read().splitlines()
    txt.insert(3, "")
    txt.insert(5, "# a comment line")
    (tmp_path / "g.in").write_text("\n".join(txt) + "\n")
    q = refio.read_flow_in(str(tmp_path / "g.in"))
    assert q["tgmin"] == p["tgmin"] and q["myorder"] == 2


def test_tecplot_file_layout_and_fields(tmp_path):
    m = M.box_mesh(3, 2, 2, jitter=0.1, seed=2)
    p = M.water_air_params(cfl=0.5)
    cu, gp = M.sphere_state(m, p, [(0.0, 0.0, 0.0, 0.3)])
    path = os.path.join(tmp_path, "tecplot0000.dat")
    refio.write_tecplot(path, m, p, cu)
    nodes, neles, blocks, conn = refio.read_tecplot(path)
    assert nodes == m.nodes and neles == m.neles and len(blocks) == 17
    assert np.array_equal(conn, m.nod.T + 1)
    assert np.allclose(blocks[0], m.x, rtol=1e-8, atol=1e-12)
    # the state was built from (p0, T0, u_in) in both phases: the recovery
    # must return them (to the file's 9 significant digits)
    ug, pg, tg, phig, phip = blocks[3], blocks[6], blocks[8], blocks[12], blocks[16]
    assert np.allclose(ug, p["ugmin"], rtol=1e-7)
    assert np.allclose(pg, p["pmin"], rtol=1e-7)
    assert np.allclose(tg, p["tgmin"], rtol=1e-7)
    assert np.allclose(phig + phip, 1.0, rtol=1e-8)
    assert phig.min() < 0.5 < phig.max()
