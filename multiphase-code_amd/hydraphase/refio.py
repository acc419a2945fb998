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


TECPLOT_VARIABLES = ('VARIABLES = "X", "Y", "Z", "U","V", "W", "PG", "PP", "TG", '
                     '"UP", "VP", "WP", "PHIG", "TP", "ROG", "ROP"')


def tecplot_fields(p: dict, cu: np.ndarray) -> list:
    """The 14 cell fields of the reference's Tecplot file, in its order
    (U V W PG PP TG UP VP WP PHIG TP ROG ROP PHIP), recovered from cu by its
    two-phase relations with clipping and cubic blending
    (main.F90:1001-1183)."""
    from .mesh import derive
    d = derive(p)
    gg, gp_, pinf, cpp = p["gamag"], p["gamap"], p["pinf"], p["cpp"]
    emin, emax, rrg, s4 = p["epslnmin"], p["epslnmax"], d["rrg"], d["foursigmabyd"]
    ug, vg, wg = cu[1] / cu[0], cu[2] / cu[0], cu[3] / cu[0]
    up, vp, wp = cu[6] / cu[5], cu[7] / cu[5], cu[8] / cu[5]
    eg = cu[4] / cu[0] - 0.5 * (ug ** 2 + vg ** 2 + wg ** 2)
    ep = cu[9] / cu[5] - 0.5 * (up ** 2 + vp ** 2 + wp ** 2)
    bbb = (s4 + pinf * gp_) - (cu[0] * (gg - 1.0) * eg) - (cu[5] * (gp_ - 1.0) * ep)
    ccc = cu[0] * (gg - 1.0) * eg * (s4 + pinf * gp_)
    pg = (-bbb + np.sqrt(bbb ** 2 + 4 * ccc)) * 0.5
    pp = pg + s4
    phig = (cu[0] * (gg - 1.0) * eg) / pg
    phip = 1.0 - phig
    rog, rop = cu[0] / phig, cu[5] / phip
    tg = pg / (rog * rrg)
    tp = ((pp + pinf) * gp_) / ((gp_ - 1.0) * rop * cpp)
    a = phig < emin
    phig = np.where(a, emin, phig); phip = np.where(a, 1.0 - phig, phip)
    ug, vg, wg, tg, pg = (np.where(a, b, c) for b, c in ((up, ug), (vp, vg), (wp, wg), (tp, tg), (pp, pg)))
    b_ = phip < emin
    phip = np.where(b_, emin, phip); phig = np.where(b_, 1.0 - phip, phig)
    up, vp, wp, tp, pp = (np.where(b_, b, c) for b, c in ((ug, up), (vg, vp), (wg, wp), (tg, tp), (pg, pp)))
    band = (phig >= emin) & (phig <= emax)
    psig = (phig - emin) / (emax - emin)
    fg = -psig * psig * (2.0 * psig - 3.0)
    ug, vg, wg, tg, pg = (np.where(band, fg * x + (1.0 - fg) * y, x)
                          for x, y in ((ug, up), (vg, vp), (wg, wp), (tg, tp), (pg, pp)))
    band = (phip >= emin) & (phip <= emax)
    psig = (phip - emin) / (emax - emin)
    fg = -psig * psig * (2.0 * psig - 3.0)
    up, vp, wp, tp, pp = (np.where(band, fg * x + (1.0 - fg) * y, x)
                          for x, y in ((up, ug), (vp, vg), (wp, wg), (tp, tg), (pp, pg)))
    rog = pg / (rrg * tg)
    rop = (pp + pinf) * gp_ / ((gp_ - 1.0) * cpp * tp)
    return [ug, vg, wg, pg, pp, tg, up, vp, wp, phig, tp, rog, rop, phip]


def write_tecplot(path: str, mesh: Mesh, p: dict, cu: np.ndarray) -> None:
    """tecplot%04d.dat as the reference writes it (main.F90:956-1266): an FE
    tetrahedral zone in block packing, node coordinates then 14 cell-centred
    fields, one ES16.8 value per line, then the 1-based connectivity.  The
    reference's header names 16 variables while 17 blocks follow (PHIP
    last); that layout is kept so tools written for its files read ours."""
    fields = tecplot_fields(p, cu)
    with open(path, "w") as fh:
        fh.write(TECPLOT_VARIABLES + "\n")
        fh.write(f"ZONE ZONETYPE=FETETRAHEDRON, NODES={mesh.nodes}, ELEMENTS={mesh.neles}, "
                 "DATAPACKING=BLOCK, VARLOCATION=([4-16]=CELLCENTERED)\n")
        for arr in [mesh.x, mesh.y, mesh.z] + fields:
            fh.write("".join(f"{v:16.8E}\n" for v in arr))
        for e in range(mesh.neles):
            n = mesh.nod[:, e] + 1
            fh.write(f" {n[0]} {n[1]} {n[2]} {n[3]}\n")


def read_tecplot(path: str):
    """(nodes, neles, blocks (17, ...), connectivity (neles, 4) 1-based)."""
    import re
    with open(path) as fh:
        fh.readline()
        zone = fh.readline()
        nodes = int(re.search(r"NODES=(\d+)", zone).group(1))
        neles = int(re.search(r"ELEMENTS=(\d+)", zone).group(1))
        blocks = [np.array([float(fh.readline()) for _ in range(nodes)]) for _ in range(3)]
        blocks += [np.array([float(fh.readline()) for _ in range(neles)]) for _ in range(14)]
        conn = np.array([[int(v) for v in fh.readline().split()] for _ in range(neles)])
    return nodes, neles, blocks, conn
