"""Tetrahedral box meshes and seeded two-phase initial states.

The reference reads its mesh from ``grid_proc%04d.dat`` / ``bc_proc%04d.in``
(sr_geometry.F90:101-139): per tet four nodes ``nod(e,1:4)``, four face
neighbours ``nc1..nc4`` (face k opposite node k), an ``iblank`` flag, and a
boundary list ``parent ghost type side`` whose ghosts are numbered
``neles+1..neles+nghosts``.  This module builds the same tables (0-based) for
an n_x x n_y x n_z hex box split into six positively oriented tets per hex
(the Kuhn split along the main diagonal), optionally jittered and with a
cylindrical hole whose faces become wall ghosts (type 5).

Boundary types follow the shipped case (bc_bg.in): x- inflow (2), x+
outflow (3), y faces (51), z faces (53).
"""
from __future__ import annotations

import dataclasses
import itertools

import numpy as np

# face k (0-based) of a tet is opposite node k; the node triple the
# reference's geometry routine uses for its area vector
# (sr_geometry.F90:149-158)
FACE_NODES = np.array([[1, 2, 3], [3, 2, 0], [1, 3, 0], [2, 1, 0]], dtype=np.int64)

BC_INFLOW, BC_OUTFLOW, BC_WALL, BC_Y, BC_Z = 2, 3, 5, 51, 53


@dataclasses.dataclass
class Mesh:
    x: np.ndarray  # (nodes,) float64
    y: np.ndarray
    z: np.ndarray
    nod: np.ndarray  # (4, neles) int32, 0-based
    nc: np.ndarray  # (4, neles) int32; >= neles names ghost neles+g
    iblank: np.ndarray  # (neles,) int32
    gparent: np.ndarray  # (nghosts,) int32
    gtype: np.ndarray  # (nghosts,) int32
    gside: np.ndarray  # (nghosts,) int32, 1..4
    shape: tuple = (0, 0, 0)

    @property
    def nodes(self) -> int:
        return int(self.x.shape[0])

    @property
    def neles(self) -> int:
        return int(self.nod.shape[1])

    @property
    def nghosts(self) -> int:
        return int(self.gparent.shape[0])

    @property
    def ntot(self) -> int:
        return self.neles + self.nghosts

    def permuted(self, perm: np.ndarray) -> "Mesh":
        """Renumber cells: new cell j is old cell perm[j].  Ghost order and
        ghost ids (neles+g) are kept; parents and neighbour ids remap."""
        ne = self.neles
        inv = np.empty(ne, dtype=np.int64)
        inv[perm] = np.arange(ne)
        nc = self.nc[:, perm].astype(np.int64)
        interior = nc < ne
        nc[interior] = inv[nc[interior]]
        return Mesh(self.x, self.y, self.z, self.nod[:, perm].copy(), nc.astype(np.int32),
                    self.iblank[perm].copy(), inv[self.gparent].astype(np.int32),
                    self.gtype.copy(), self.gside.copy(), self.shape)


def _kuhn_tets():
    """Six tets of the unit cube along the 0-7 diagonal, as corner ids
    (corner c = i + 2j + 4k), each positively oriented."""
    corners = np.array([[i & 1, (i >> 1) & 1, (i >> 2) & 1] for i in range(8)], dtype=np.float64)
    tets = []
    for a, b, _ in itertools.permutations(range(3)):
        c1 = 1 << a
        c2 = c1 | (1 << b)
        t = [0, c1, c2, 7]
        p = corners[t]
        v = np.linalg.det(np.stack([p[1] - p[0], p[2] - p[0], p[3] - p[0]]))
        if v < 0:
            t[2], t[3] = t[3], t[2]
        tets.append(t)
    return np.array(tets, dtype=np.int64)


def box_mesh(nx: int, ny: int, nz: int, *, lo=(-0.5, -0.5, -0.5), hi=(0.5, 0.5, 0.5),
             jitter: float = 0.1, seed: int = 1, hole=None) -> Mesh:
    """Kuhn-split hex box.  ``jitter`` moves interior nodes uniformly by
    +-jitter*h per axis (seeded).  ``hole=(xc, yc, r)`` removes the tets whose
    centroid lies inside the z-aligned cylinder; the faces they shared become
    wall ghosts (type 5)."""
    x, y, z, nod = _box_nodes_tets(nx, ny, nz, lo, hi, jitter, seed)
    if hole is not None:
        xc, yc, r = hole
        cx = x[nod].mean(axis=1)
        cy = y[nod].mean(axis=1)
        keep = (cx - xc) ** 2 + (cy - yc) ** 2 > r * r
        nod = nod[keep]
    return _connect(x, y, z, nod, np.asarray(lo, dtype=np.float64), np.asarray(hi, dtype=np.float64),
                    (nx, ny, nz))


def _box_nodes_tets(nx, ny, nz, lo, hi, jitter, seed):
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    h = (hi - lo) / np.array([nx, ny, nz], dtype=np.float64)
    ii, jj, kk = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    ii = ii.transpose(2, 1, 0).ravel()
    jj = jj.transpose(2, 1, 0).ravel()
    kk = kk.transpose(2, 1, 0).ravel()
    x = lo[0] + ii * h[0]
    y = lo[1] + jj * h[1]
    z = lo[2] + kk * h[2]
    if jitter:
        rng = np.random.default_rng(seed)
        d = rng.uniform(-jitter, jitter, size=(3, x.size))
        inner = (ii > 0) & (ii < nx) & (jj > 0) & (jj < ny) & (kk > 0) & (kk < nz)
        x = np.where(inner, x + d[0] * h[0], x)
        y = np.where(inner, y + d[1] * h[1], y)
        z = np.where(inner, z + d[2] * h[2], z)

    def nid(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    hi_, hj_, hk_ = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    hi_ = hi_.transpose(2, 1, 0).ravel()
    hj_ = hj_.transpose(2, 1, 0).ravel()
    hk_ = hk_.transpose(2, 1, 0).ravel()
    corner_nodes = np.stack([nid(hi_ + (c & 1), hj_ + ((c >> 1) & 1), hk_ + ((c >> 2) & 1))
                             for c in range(8)], axis=1)  # (nhex, 8)
    kt = _kuhn_tets()
    nod = corner_nodes[:, kt].reshape(-1, 4)  # (neles, 4), hex-major
    return x, y, z, nod


def annulus_mesh(nr: int, ntheta: int, nz: int, *, r0: float = 0.1, r1: float = 0.25, center=(0.0, 0.0),
                 zlo: float = -0.5, zhi: float = 0.5, box_lo=(-0.5, -0.5, -0.5), box_hi=(0.5, 0.5, 0.5)):
    """Near-body mesh around a z-aligned cylinder: the shell r0 <= r <= r1
    split into nr x ntheta x nz curved hexes (periodic in theta), each
    Kuhn-split into six tets in (r, theta, z) index space, which maps
    (r, theta, z) -> (x, y, z) with positive orientation.  Returns
    (x, y, z, nod (ne, 4)) for combine(); the inner surface becomes a wall
    (type 5), z faces type 53; the outer surface is left to the overset
    fringe (type 5 ghosts on cells that are receptors)."""
    i, j, k = np.meshgrid(np.arange(nr + 1), np.arange(ntheta), np.arange(nz + 1), indexing="ij")
    i, j, k = i.transpose(2, 1, 0).ravel(), j.transpose(2, 1, 0).ravel(), k.transpose(2, 1, 0).ravel()
    r = r0 + (r1 - r0) * i / nr
    th = 2.0 * np.pi * j / ntheta
    x = center[0] + r * np.cos(th)
    y = center[1] + r * np.sin(th)
    z = zlo + (zhi - zlo) * k / nz

    def nid(a, b, c):
        return a + (nr + 1) * ((b % ntheta) + ntheta * c)

    ha, hb, hc = np.meshgrid(np.arange(nr), np.arange(ntheta), np.arange(nz), indexing="ij")
    ha, hb, hc = ha.transpose(2, 1, 0).ravel(), hb.transpose(2, 1, 0).ravel(), hc.transpose(2, 1, 0).ravel()
    corner = np.stack([nid(ha + (c & 1), hb + ((c >> 1) & 1), hc + ((c >> 2) & 1)) for c in range(8)], axis=1)
    nod = corner[:, _kuhn_tets()].reshape(-1, 4)
    return x, y, z, nod


def combine(parts, lo=(-0.5, -0.5, -0.5), hi=(0.5, 0.5, 0.5), shape=(0, 0, 0)) -> Mesh:
    """One mesh of several disconnected node/tet sets [(x, y, z, nod), ...]
    (cells numbered part by part); faces and boundary types as box_mesh."""
    xs, ys, zs, nods, off = [], [], [], [], 0
    for (x, y, z, nod) in parts:
        xs.append(x); ys.append(y); zs.append(z)
        nods.append(np.asarray(nod, dtype=np.int64) + off)
        off += x.size
    return _connect(np.concatenate(xs), np.concatenate(ys), np.concatenate(zs), np.concatenate(nods),
                    np.asarray(lo, dtype=np.float64), np.asarray(hi, dtype=np.float64), shape)


def box_tets(nx: int, ny: int, nz: int, *, lo=(-0.5, -0.5, -0.5), hi=(0.5, 0.5, 0.5), jitter: float = 0.1,
             seed: int = 1, keep=None):
    """(x, y, z, nod (ne, 4)) of the Kuhn-split box of box_mesh; keep(cx, cy,
    cz) -> bool mask selects the tets to keep by centroid."""
    x, y, z, nod = _box_nodes_tets(nx, ny, nz, lo, hi, jitter, seed)
    if keep is not None:
        c = [a[nod].mean(axis=1) for a in (x, y, z)]
        nod = nod[keep(*c)]
    return x, y, z, nod


def _connect(x, y, z, nod, lo, hi, shape) -> Mesh:
    ne = nod.shape[0]
    fn = nod[:, FACE_NODES]  # (ne, 4, 3)
    key_nodes = np.sort(fn, axis=2).reshape(-1, 3)
    nn = np.int64(x.size)
    key = (key_nodes[:, 0] * nn + key_nodes[:, 1]) * nn + key_nodes[:, 2]
    order = np.argsort(key, kind="stable")
    ks = key[order]
    same_next = np.zeros(ks.size, dtype=bool)
    same_next[:-1] = ks[1:] == ks[:-1]
    nc = np.full(ne * 4, -1, dtype=np.int64)
    a = order[:-1][same_next[:-1]]
    b = order[1:][same_next[:-1]]
    nc[a] = b // 4
    nc[b] = a // 4
    bnd = np.nonzero(nc < 0)[0]
    cells = bnd // 4
    sides = bnd % 4
    tri = fn.reshape(-1, 3)[bnd]
    tol = 1e-12
    btype = np.full(bnd.size, BC_WALL, dtype=np.int64)
    on = lambda arr, v: np.all(np.abs(arr[tri] - v) < tol, axis=1)
    btype[on(y, lo[1]) | on(y, hi[1])] = BC_Y
    btype[on(z, lo[2]) | on(z, hi[2])] = BC_Z
    btype[on(x, lo[0])] = BC_INFLOW
    btype[on(x, hi[0])] = BC_OUTFLOW
    # ghost order: by parent cell, then side (the reference takes the list
    # order of bc_proc%04d.in as given)
    g_order = np.lexsort((sides, cells))
    cells, sides, btype, bnd = cells[g_order], sides[g_order], btype[g_order], bnd[g_order]
    nc[bnd] = ne + np.arange(bnd.size)
    nc = nc.reshape(ne, 4).T.astype(np.int32)
    return Mesh(x=np.ascontiguousarray(x), y=np.ascontiguousarray(y), z=np.ascontiguousarray(z),
                nod=np.ascontiguousarray(nod.T.astype(np.int32)), nc=np.ascontiguousarray(nc),
                iblank=np.ones(ne, dtype=np.int32), gparent=cells.astype(np.int32),
                gtype=btype.astype(np.int32), gside=(sides + 1).astype(np.int32), shape=tuple(shape))


def morton_order(mesh: Mesh, bits: int = 10) -> np.ndarray:
    """Cell permutation sorting centroids along a Z-order curve, so that
    face neighbours sit close in memory."""
    cx = mesh.x[mesh.nod].mean(axis=0)
    cy = mesh.y[mesh.nod].mean(axis=0)
    cz = mesh.z[mesh.nod].mean(axis=0)

    def q(v):
        v = (v - v.min()) / max(v.max() - v.min(), 1e-300)
        return np.minimum((v * (1 << bits)).astype(np.uint64), (1 << bits) - 1)

    def spread(v):
        v = v.astype(np.uint64)
        out = np.zeros_like(v)
        for b in range(bits):
            out |= ((v >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b)
        return out

    code = spread(q(cx)) | (spread(q(cy)) << np.uint64(1)) | (spread(q(cz)) << np.uint64(2))
    return np.argsort(code, kind="stable")


def slab_order(mesh: Mesh, ny: int, nz: int, *, by: int = 8, bz: int = 8, lo=(-0.5, -0.5, -0.5),
               hi=(0.5, 0.5, 0.5), cells=None) -> np.ndarray:
    """Cell permutation for gather locality on a mesh generated from an
    ny x nz (in y, z) structured box with x fastest (box_mesh, box_tets):
    cells sorted by (z slab of bz layers, y slab of by rows, original
    index).  A y or z face neighbour then sits about nx*by or nx*by*bz
    hexes away instead of nx or nx*ny, except across slab boundaries.
    ``cells=(start, stop)`` reorders only that range (another mesh's cells
    appended after it keep their place).  Per-cell results do not depend on
    the numbering; the permutation is for Mesh.permuted (new cell j is old
    cell perm[j])."""
    ne = mesh.neles
    start, stop = (0, ne) if cells is None else cells
    cy = mesh.y[mesh.nod].mean(axis=0)
    cz = mesh.z[mesh.nod].mean(axis=0)
    iy = np.clip(np.floor((cy - lo[1]) / (hi[1] - lo[1]) * ny), 0, ny - 1).astype(np.int64)
    iz = np.clip(np.floor((cz - lo[2]) / (hi[2] - lo[2]) * nz), 0, nz - 1).astype(np.int64)
    idx = np.arange(ne)
    inside = (idx >= start) & (idx < stop)
    # the reordered range keeps its place among the others
    seg = np.where(idx < start, 0, np.where(inside, 1, 2))
    zb = np.where(inside, iz // bz, 0)
    yb = np.where(inside, iy // by, 0)
    return np.lexsort((idx, yb, zb, seg))


# --------------------------------------------------------------------------
# flow parameters and initial states
# --------------------------------------------------------------------------

PARAM_FIELDS = [
    "pmax", "pmin", "tgmax", "tgmin", "ugmax", "ugmin", "vgmax", "vgmin", "wgmax", "wgmin",
    "upmax", "upmin", "vpmax", "vpmin", "wpmax", "wpmin", "tpmax", "tpmin",
    "pinf", "gamap", "gamag", "cpp", "cpg", "amolg", "prg",
    "epslnmax", "epslnmin", "epslninf", "phipzero",
    "cfl", "diap", "surfacetension", "as_", "ts",
]
PARAM_INT_FIELDS = ["niter", "itersub", "istart", "myflux", "myorder", "pad_"]
PARAM_DERIVED = ["co2", "co4", "c23", "cmug", "rrg", "gravity", "prp", "foursigmabyd"]


def water_air_params(*, mach: float = 0.3, cfl: float = 0.5, viscous_gas: bool = False,
                     myflux: int = 2, p0: float = 1.0e5, t0: float = 300.0,
                     mach_phase: str = "gas") -> dict:
    """Air (gamma 1.4) carrying water (stiffened gas, gamma 4.4, p_inf 6e8);
    the inflow carries liquid fraction phipzero = 1e-6.
    The inflow speed is ``mach`` times the sound speed of ``mach_phase``
    ("gas": air at (p0, t0); "liquid": water at (p0, t0)).
    cpp is chosen so that the water density is 1000 kg/m^3 at (p0, t0).
    The gas viscosity is Sutherland with 1.8e-5 Pa s at t0 when
    ``viscous_gas`` (as=0 otherwise); the liquid viscosity is the reference's
    hard-wired 1e-3 Pa s (sr_viscousgp.F90:563)."""
    gamap, pinf = 4.4, 6.0e8
    rho_w = 1000.0
    cpp = (p0 + pinf) * gamap / ((gamap - 1.0) * rho_w * t0)
    a_w = np.sqrt(gamap * (p0 + pinf) / rho_w)
    a_g = np.sqrt(1.4 * 287.0 * t0)
    u = float(mach * (a_g if mach_phase == "gas" else a_w))
    ts = 110.4
    as_ = 1.8e-5 * (1.0 + ts / t0) / np.sqrt(t0) if viscous_gas else 0.0
    return dict(
        pmax=p0, pmin=p0, tgmax=t0, tgmin=t0, ugmax=u, ugmin=u, vgmax=0.0, vgmin=0.0,
        wgmax=0.0, wgmin=0.0, upmax=u, upmin=u, vpmax=0.0, vpmin=0.0, wpmax=0.0, wpmin=0.0,
        tpmax=t0, tpmin=t0, pinf=pinf, gamap=gamap, gamag=1.4, cpp=float(cpp), cpg=1004.5,
        amolg=28.9, prg=0.72, epslnmax=1.0e-3, epslnmin=1.0e-8, epslninf=2.0e-8,
        phipzero=1.0e-6, cfl=cfl, diap=100.0e-6, surfacetension=0.0, as_=float(as_), ts=ts,
        niter=1, itersub=100, istart=1, myflux=myflux, myorder=2, pad_=0,
    )


def derive(p: dict) -> dict:
    """main.F90:336-349 (REAL*4 literals as the Fortran evaluates them)."""
    f32 = np.float32
    d = dict(p)
    d["co2"] = float(f32(1.0) / f32(16.0))
    d["co4"] = float(f32(1.0) / f32(1024.0))
    d["c23"] = float(f32(2.0) / f32(3.0))
    d["cmug"] = float(np.float64(f32(11.848e-8)) * np.sqrt(p["amolg"]))
    d["rrg"] = 287.0
    d["gravity"] = 0.0
    d["prp"] = 10.0
    d["foursigmabyd"] = 4.0 * p["surfacetension"] / p["diap"]
    return d


def prims_to_conserved(p: dict, phig, pg, tg, ug, vg, wg, tp, up, vp, wp):
    """Conserved variables from primitives through the reference's forward
    relations (sr_primitive.F90:144-161 with pp = pg + 4 sigma/d)."""
    d = derive(p)
    rrg, gamag, gamap, pinf, cpp = d["rrg"], p["gamag"], p["gamap"], p["pinf"], p["cpp"]
    pp = pg + d["foursigmabyd"]
    phip = 1.0 - phig
    rog = pg / (rrg * tg)
    rop = (pp + pinf) * gamap / ((gamap - 1.0) * cpp * tp)
    eg = pg / ((gamag - 1.0) * rog)
    ep = (pp + gamap * pinf) / ((gamap - 1.0) * rop)
    egt = eg + 0.5 * (ug * ug + vg * vg + wg * wg)
    ept = ep + 0.5 * (up * up + vp * vp + wp * wp)
    cu = np.stack([phig * rog, phig * rog * ug, phig * rog * vg, phig * rog * wg, phig * rog * egt,
                   phip * rop, phip * rop * up, phip * rop * vp, phip * rop * wp, phip * rop * ept])
    return cu, rog


def sphere_state(mesh: Mesh, p: dict, spheres, *, width_cells: float = 2.0):
    """Restart state: uniform (p0, T0, u_in) air carrying water spheres.
    Each sphere (xc, yc, zc, R) contributes a liquid fraction
    alpha = 0.5(1 - tanh((r-R)/w)) with w = width_cells*h/2 (a 2-cell-wide
    interface for width_cells=2); the gas fraction 1-alpha is clipped to
    [1e-6, 1-1e-6].  Air is the carrier phase as in the shipped case
    (flow.in: epslninf=2e-8 liquid background, sr_init_cond.F90:177-195).
    Returns (cu (10, neles), ghost gas primitives (6, nghosts): ug vg wg pg
    tg rog) in the layout of restart%04d.in (main.F90:431-438)."""
    cx = mesh.x[mesh.nod].mean(axis=0)
    cy = mesh.y[mesh.nod].mean(axis=0)
    cz = mesh.z[mesh.nod].mean(axis=0)
    nx, ny, nz = mesh.shape
    span = max(mesh.x.max() - mesh.x.min(), 1e-300)
    h = span / max(nx, 1)
    w = 0.5 * width_cells * h
    alpha = np.zeros(mesh.neles)
    for (bx, by, bz, r) in spheres:
        rr = np.sqrt((cx - bx) ** 2 + (cy - by) ** 2 + (cz - bz) ** 2)
        alpha = np.maximum(alpha, 0.5 * (1.0 - np.tanh((rr - r) / w)))
    phig = np.clip(1.0 - alpha, 1.0e-6, 1.0 - 1.0e-6)
    ne = mesh.neles
    one = np.ones(ne)
    u = p["ugmin"]
    cu, rog = prims_to_conserved(p, phig, p["pmin"] * one, p["tgmin"] * one, u * one, 0 * one, 0 * one,
                                 p["tpmin"] * one, u * one, 0 * one, 0 * one)
    gpar = mesh.gparent
    ngh = mesh.nghosts
    gp = np.stack([u * np.ones(ngh), np.zeros(ngh), np.zeros(ngh),
                   p["pmin"] * np.ones(ngh), p["tgmin"] * np.ones(ngh), rog[gpar]])
    return np.ascontiguousarray(cu), np.ascontiguousarray(gp)


def smooth_state(mesh: Mesh, p: dict, *, amp: float = 0.05, center=(0.0, 0.0, 0.0), sigma: float = 0.15):
    """Restart state without an interface: the air carrier everywhere (gas
    fraction 1 - phipzero, the inflow's liquid trace), the same velocity and
    temperature in both phases, and a smooth Gaussian bump of relative
    height ``amp`` and width ``sigma`` on the pressure and the x velocity.
    Returns (cu, ghost gas primitives) like sphere_state."""
    cx = mesh.x[mesh.nod].mean(axis=0)
    cy = mesh.y[mesh.nod].mean(axis=0)
    cz = mesh.z[mesh.nod].mean(axis=0)
    r2 = (cx - center[0]) ** 2 + (cy - center[1]) ** 2 + (cz - center[2]) ** 2
    bump = amp * np.exp(-r2 / (sigma * sigma))
    ne = mesh.neles
    one = np.ones(ne)
    u = p["ugmin"] * (1.0 + bump)
    phig = (1.0 - p["phipzero"]) * one
    cu, rog = prims_to_conserved(p, phig, p["pmin"] * (1.0 + bump), p["tgmin"] * one, u, 0 * one, 0 * one,
                                 p["tpmin"] * one, u, 0 * one, 0 * one)
    ngh = mesh.nghosts
    gp = np.stack([p["ugmin"] * np.ones(ngh), np.zeros(ngh), np.zeros(ngh),
                   p["pmin"] * np.ones(ngh), p["tgmin"] * np.ones(ngh), rog[mesh.gparent]])
    return np.ascontiguousarray(cu), np.ascontiguousarray(gp)


def sphere_train(seed: int, n: int = 8, rmin: float = 0.03, rmax: float = 0.08,
                 lo=(-0.5, -0.5, -0.5), hi=(0.5, 0.5, 0.5), avoid=None):
    """n spheres with radii U[rmin, rmax], clear of each other and of
    ``avoid`` (a (xc, yc, r) z-cylinder), seeded."""
    rng = np.random.default_rng(seed)
    out = []
    lo = np.asarray(lo)
    hi = np.asarray(hi)
    while len(out) < n:
        r = rng.uniform(rmin, rmax)
        c = rng.uniform(lo + r + 0.02, hi - r - 0.02)
        if avoid is not None and np.hypot(c[0] - avoid[0], c[1] - avoid[1]) < avoid[2] + r + 0.02:
            continue
        if any(np.linalg.norm(c - np.asarray(b[:3])) < r + b[3] for b in out):
            continue
        out.append((float(c[0]), float(c[1]), float(c[2]), float(r)))
    return out
