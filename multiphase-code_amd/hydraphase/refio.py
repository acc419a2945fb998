"""Readers and writers for the reference's own input/output files.

  geometry_files/grid_proc%04d.dat  nodes neles; n x y z g_id tag;
                                    e n1..n4 nc1..nc4 iblank  (sr_geometry.F90:101-128)
  geometry_files/bc_proc%04d.in     nghosts; parent ghost type side (:131-134)
  geometry_files/proc_bc%04d.in     pghosts max; 8 ints per row (:136-139)
  flow.in                           positional tuples, '#' comments (sr_Read_flow.F90:19-268)
  restart_files/restart%04d.in      time; cu(neles,neq); ghost ug vg wg pg tg rog
                                    (main.F90:427-441, :812-824)
  itter.in                          one integer (main.F90:439, :827)

Ids in the files are 1-based; the Mesh tables are 0-based.
"""
from __future__ import annotations

import os

import numpy as np

from .mesh import Mesh

# (fields, comment) of each flow.in tuple in file order
FLOW_TUPLES = [
    (("pmax", "pmin"), "pmax, pmin"),
    (("tgmax", "tgmin"), "tgmax, tgmin"),
    (("ugmax", "ugmin"), "ugmax, ugmin"),
    (("vgmax", "vgmin"), "vgmax, vgmin"),
    (("wgmax", "wgmin"), "wgmax, wgmin"),
    (("upmax", "upmin"), "upmax, upmin"),
    (("vpmax", "vpmin"), "vpmax, vpmin"),
    (("wpmax", "wpmin"), "wpmax, wpmin"),
    (("tpmax", "tpmin"), "tpmax, tpmin"),
    (("pinf", "gamap"), "pinf, gamap"),
    (("gamag", "cpp"), "gamag, cpp"),
    (("cpg", "amolg"), "cpg, amolg"),
    (("prg", "epslnmax"), "prg, epslnmax"),
    (("epslnmin", "epslninf", "phipzero"), "epslnmin, epslninf, phipzero"),
    (("cfl", "niter", "itersub"), "cfl, niter, itersub"),
    (("istart", "diap"), "istart, diap"),
    (("surfacetension", "myflux"), "surfacetension, myflux"),
    (("myorder", "as_", "ts"), "myorder, as, ts"),
    (("rhoq", "rhor", "rhos"), "rhoq, rhor, rhos (five-phase, unused)"),
    (("radq", "radr", "rads"), "radq, radr, rads (five-phase, unused)"),
    (("phiginit", "phipinit", "phiqinit", "phirinit", "phisinit"), "initial fractions (five-phase)"),
    (("cp_t", "cp_o", "cp_p", "amol_f", "amol_o", "amol_p"), "species (unused)"),
    (("tolsp", "sratio", "hfp", "preexp", "ebyr"), "reaction (unused)"),
]
INT_FIELDS = {"niter", "itersub", "istart", "myflux", "myorder"}


def _f(v: float) -> str:
    return repr(float(v))


def write_flow_in(path: str, p: dict) -> None:
    lines = ["# flow.in written by hydraphase.refio (field order of sr_Read_flow.F90)"]
    for fields, comment in FLOW_TUPLES:
        vals = []
        for f in fields:
            v = p.get(f, 0)
            vals.append(str(int(v)) if f in INT_FIELDS else _f(v))
        lines.append("   ".join(vals) + "   # " + comment)
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def read_flow_in(path: str) -> dict:
    """Same rules as sr_Read_flow.F90: skip blank and '#' lines, drop a
    trailing comment, take the next line that parses as the wanted tuple."""
    with open(path) as fh:
        raw = fh.read().splitlines()
    out, it = {}, iter(raw)
    for fields, _ in FLOW_TUPLES:
        for line in it:
            if not line.strip() or line.lstrip().startswith("#") or line.startswith("#"):
                continue
            line = line.split("#", 1)[0]
            toks = line.replace(",", " ").split()
            if len(toks) < len(fields):
                continue
            try:
                vals = [int(float(t)) if f in INT_FIELDS else float(t.replace("d", "e").replace("D", "e"))
                        for f, t in zip(fields, toks)]
            except ValueError:
                continue
            out.update(dict(zip(fields, vals)))
            break
    return out


def write_mesh(dirname: str, mesh: Mesh, rank: int = 0) -> None:
    os.makedirs(dirname, exist_ok=True)
    ne, ng = mesh.neles, mesh.nghosts
    with open(os.path.join(dirname, f"grid_proc{rank:04d}.dat"), "w") as fh:
        fh.write(f"{mesh.nodes} {ne}\n")
        ids = np.arange(1, mesh.nodes + 1)
        np.savetxt(fh, np.column_stack([ids, mesh.x, mesh.y, mesh.z, ids, np.ones(mesh.nodes)]),
                   fmt=["%d", "%.17g", "%.17g", "%.17g", "%d", "%d"])
        np.savetxt(fh, np.column_stack([np.arange(1, ne + 1), (mesh.nod + 1).T, (mesh.nc + 1).T, mesh.iblank]),
                   fmt="%d")
    with open(os.path.join(dirname, f"bc_proc{rank:04d}.in"), "w") as fh:
        fh.write(f"{ng}\n")
        np.savetxt(fh, np.column_stack([mesh.gparent + 1, np.arange(ne + 1, ne + ng + 1), mesh.gtype, mesh.gside]),
                   fmt="%d")
    with open(os.path.join(dirname, f"proc_bc{rank:04d}.in"), "w") as fh:
        fh.write("0 0\n")


def read_mesh(dirname: str, rank: int = 0) -> Mesh:
    with open(os.path.join(dirname, f"grid_proc{rank:04d}.dat")) as fh:
        nodes, ne = (int(t) for t in fh.readline().split()[:2])
        nd = np.loadtxt(fh, max_rows=nodes, usecols=(0, 1, 2, 3))
        el = np.loadtxt(fh, max_rows=ne, dtype=np.int64)
    x = np.empty(nodes); y = np.empty(nodes); z = np.empty(nodes)
    idx = nd[:, 0].astype(np.int64) - 1
    x[idx], y[idx], z[idx] = nd[:, 1], nd[:, 2], nd[:, 3]
    el = el[np.argsort(el[:, 0])]
    bc = np.loadtxt(os.path.join(dirname, f"bc_proc{rank:04d}.in"), skiprows=1, dtype=np.int64, ndmin=2)
    return Mesh(x=x, y=y, z=z, nod=np.ascontiguousarray((el[:, 1:5] - 1).T.astype(np.int32)),
                nc=np.ascontiguousarray((el[:, 5:9] - 1).T.astype(np.int32)),
                iblank=el[:, 9].astype(np.int32), gparent=(bc[:, 0] - 1).astype(np.int32),
                gtype=bc[:, 2].astype(np.int32), gside=bc[:, 3].astype(np.int32))


def write_restart(path: str, time: float, cu: np.ndarray, gp: np.ndarray) -> None:
    """cu (10, neles), gp (6, nghosts); one record per line as the
    reference's list-directed writes (values round-trip exactly)."""
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    with open(path, "w") as fh:
        fh.write(_f(time) + "\n")
        fh.write(" ".join(_f(v) for v in np.asarray(cu).ravel()) + "\n")
        for k in range(6):
            fh.write(" ".join(_f(v) for v in np.asarray(gp[k])) + "\n")


def read_restart(path: str, neles: int, nghosts: int):
    vals = np.array(open(path).read().replace("D", "E").replace("d", "e").split(), dtype=np.float64)
    t = vals[0]
    cu = vals[1:1 + 10 * neles].reshape(10, neles)
    gp = vals[1 + 10 * neles:1 + 10 * neles + 6 * nghosts].reshape(6, nghosts)
    return t, cu, gp


def write_case(dirname: str, mesh: Mesh, p: dict, cu, gp, time: float = 0.0, itter: int = 0) -> None:
    """A complete run directory the reference (and the drop-in driver) can
    start from with istart=1."""
    write_mesh(os.path.join(dirname, "geometry_files"), mesh)
    q = dict(p)
    q["istart"] = 1
    write_flow_in(os.path.join(dirname, "flow.in"), q)
    write_restart(os.path.join(dirname, "restart_files", "restart0000.in"), time, cu, gp)
    with open(os.path.join(dirname, "itter.in"), "w") as fh:
        fh.write(f"{itter}\n")
