"""Overset coupling by fixed barycentric donor lists.

The reference has no donor interpolation of its own (only iblank freezing;
its overset preprocessing is an external library that is not part of it).
Here a receptor cell r takes its conserved variables as a fixed
combination of four donor cells of the other mesh,

    cn(r, v) = w1 cn(d1, v) + w2 cn(d2, v) + w3 cn(d3, v) + w4 cn(d4, v),

summed in that order, after every RK stage (libhydra_hip.so, hp_set_donors).
The donors are four active cells of the other mesh whose centroids span a
tetrahedron containing the receptor's centroid, and w are the receptor's
barycentric coordinates in it: the weights are non-negative, sum to one and
reproduce any linear field exactly.  The few receptors whose centroid lies
outside the hull of the other mesh's donor centroids (in its first or last
cell layer along a boundary it shares) have no containing tetrahedron;
they take the best one's weights with the negative ones dropped, a convex
combination.  A receptor whose centroid coincides with a candidate's takes
that cell alone (weights 1, 0, 0, 0).

File format (ours; the reference has none): ``donor_proc%04d.in``, first
line the receptor count, then one line per receptor
``receptor d1 d2 d3 d4 w1 w2 w3 w4`` with 1-based cell ids.
"""
from __future__ import annotations

import itertools
import os

import numpy as np

from . import mesh as M

COMBOS = np.array(list(itertools.combinations(range(8), 4)), dtype=np.int64)  # 70 x 4


def centroids(m: M.Mesh):
    return np.stack([m.x[m.nod].mean(axis=0), m.y[m.nod].mean(axis=0), m.z[m.nod].mean(axis=0)], axis=1)


def barycentric(p, a, b, c, d):
    """Barycentric coordinates of points p (n,3) in tets (a,b,c,d) (n,3)
    each, by Cramer's rule; returns (n, 4) and the tet volumes x6 (n,)."""
    v0, v1, v2, dp = b - a, c - a, d - a, p - a
    c12 = np.cross(v1, v2)
    det = np.einsum("ij,ij->i", v0, c12)
    inv = np.divide(1.0, det, out=np.zeros_like(det), where=np.abs(det) > 1e-300)
    l1 = np.einsum("ij,ij->i", dp, c12) * inv
    l2 = np.einsum("ij,ij->i", v0, np.cross(dp, v2)) * inv
    l3 = np.einsum("ij,ij->i", v0, np.cross(v1, dp)) * inv
    w = np.stack([1.0 - l1 - l2 - l3, l1, l2, l3], axis=1)
    return w, det


def _best_tets(points, cand_xyz, cand_ids, k):
    """For each point, among all 4-subsets of its k nearest candidate
    centroids, the non-degenerate one whose smallest barycentric weight is
    largest: (donor ids, weights, that smallest weight, nearest distance,
    local cell scale)."""
    from scipy.spatial import cKDTree
    import itertools
    tree = cKDTree(cand_xyz)
    dist, nn = tree.query(points, k=k)
    n = points.shape[0]
    best_w = np.full((n, 4), np.nan)
    best_d = np.zeros((n, 4), dtype=np.int64)
    best_min = np.full(n, -np.inf)
    scale = np.linalg.norm(cand_xyz[nn[:, 1]] - cand_xyz[nn[:, 0]], axis=1) ** 3
    combos = COMBOS if k == 8 else np.array(list(itertools.combinations(range(k), 4)), dtype=np.int64)
    for combo in combos:
        idx = nn[:, combo]  # (n, 4)
        w, det = barycentric(points, *(cand_xyz[idx[:, q]] for q in range(4)))
        # reject flat tets: |det| (6 x volume) well below the local cell scale
        good = np.abs(det) > 1e-3 * scale
        mn = np.where(good, w.min(axis=1), -np.inf)
        better = mn > best_min
        best_min = np.where(better, mn, best_min)
        best_w[better] = w[better]
        best_d[better] = cand_ids[idx[better]]
    return best_d, best_w, best_min, dist[:, 0], scale, nn


def find_donors(points: np.ndarray, cand_xyz: np.ndarray, cand_ids: np.ndarray, k: int = 8,
                widen=(16, 24), tol: float = 1e-12):
    """For each point, four donors whose centroids span a tetrahedron that
    contains it, and its barycentric weights (all >= -tol).  The search
    takes the best tetrahedron over the k nearest candidate centroids, then
    over the `widen` nearest for the points still outside theirs.  A point
    outside every candidate tetrahedron (beyond the hull of the candidate
    centroids, e.g. in a boundary layer of the other mesh) takes the best
    one's weights with the negative ones set to zero and the rest scaled to
    sum to one: a convex combination, bounded by its donors' values though
    no longer exact for linear fields.  Returns (donors, weights, the
    smallest weight before that clipping)."""
    best_d, best_w, best_min, d0, scale, nn = _best_tets(points, cand_xyz, cand_ids, k)
    for kk in widen:
        out = np.nonzero(best_min < -tol)[0]
        if out.size == 0 or kk > cand_xyz.shape[0]:
            break
        d2, w2, m2, _, _, _ = _best_tets(points[out], cand_xyz, cand_ids, kk)
        better = m2 > best_min[out]
        sel = out[better]
        best_d[sel], best_w[sel], best_min[sel] = d2[better], w2[better], m2[better]
    # a receptor on a candidate centroid takes that cell alone
    on = d0 <= 1e-12 * np.cbrt(scale)
    if on.any():
        best_d[on] = cand_ids[nn[on, :4]]
        best_w[on] = np.array([1.0, 0.0, 0.0, 0.0])
        best_min[on] = 0.0
    if not np.isfinite(best_min).all():
        raise ValueError("donor search: some receptors have no non-degenerate donor tetrahedron")
    clip = best_min < -tol
    if clip.any():
        w = np.maximum(best_w[clip], 0.0)
        best_w[clip] = w / w.sum(axis=1, keepdims=True)
    return best_d, best_w, best_min


def write_donors(path: str, recv: np.ndarray, donors: np.ndarray, weights: np.ndarray):
    """donor_proc%04d.in: count, then receptor d1..d4 w1..w4 (1-based ids)."""
    with open(path, "w") as fh:
        fh.write(f"{recv.size}\n")
        for r, d, w in zip(recv, donors, weights):
            fh.write(f"{r + 1} {d[0] + 1} {d[1] + 1} {d[2] + 1} {d[3] + 1} "
                     f"{w[0]:.17e} {w[1]:.17e} {w[2]:.17e} {w[3]:.17e}\n")


def read_donors(path: str):
    with open(path) as fh:
        n = int(fh.readline())
        rows = [fh.readline().split() for _ in range(n)]
    recv = np.array([int(r[0]) - 1 for r in rows], dtype=np.int64)
    donors = np.array([[int(v) - 1 for v in r[1:5]] for r in rows], dtype=np.int64).reshape(-1, 4)
    weights = np.array([[float(v) for v in r[5:9]] for r in rows]).reshape(-1, 4)
    return recv, donors, weights


def overset_mesh(nx: int = 96, ny: int = 96, nz: int = 90, *, nr: int = 10, ntheta: int = 112,
                 r_body: float = 0.1, r_outer: float = 0.25, r_cut: float = 0.15, fringe: int = 2,
                 jitter: float = 0.1, seed: int = 1, slabs=None):
    """Background box (Kuhn-split, jittered) with the cells of centroid
    radius < r_cut removed, plus the near-body annulus r_body..r_outer
    (nr x ntheta x nz hexes x 6 tets), in one mesh: background cells first.
    Receptors (iblank 0, filled from the other mesh) are the background
    cells within `fringe` background spacings outside r_cut and the annulus
    cells of the outer `fringe` radial layers.  ``slabs=(by, bz)`` numbers
    the background cells in y-z slabs (mesh.slab_order) for gather
    locality; the annulus keeps its order.  Returns (mesh, recv, donors,
    weights) with 0-based ids."""
    bg = M.box_tets(nx, ny, nz, jitter=jitter, seed=seed,
                    keep=lambda cx, cy, cz: cx * cx + cy * cy >= r_cut * r_cut)
    an = M.annulus_mesh(nr, ntheta, nz, r0=r_body, r1=r_outer)
    m = M.combine([bg, an], shape=(nx, ny, nz))
    nbg = bg[3].shape[0]
    if slabs is not None:
        m = m.permuted(M.slab_order(m, ny, nz, by=slabs[0], bz=slabs[1], cells=(0, nbg)))
    c = centroids(m)
    rad = np.hypot(c[:, 0], c[:, 1])
    h = 1.0 / nx
    dr = (r_outer - r_body) / nr
    is_bg = np.arange(m.neles) < nbg
    recv_bg = np.nonzero(is_bg & (rad < r_cut + fringe * h))[0]
    recv_an = np.nonzero(~is_bg & (rad > r_outer - fringe * dr))[0]
    ib = np.ones(m.neles, dtype=np.int32)
    ib[recv_bg] = 0
    ib[recv_an] = 0
    m.iblank = ib
    act_bg = np.nonzero(is_bg & (ib == 1))[0]
    act_an = np.nonzero(~is_bg & (ib == 1))[0]
    d1, w1, _ = find_donors(c[recv_bg], c[act_an], act_an)
    d2, w2, _ = find_donors(c[recv_an], c[act_bg], act_bg)
    recv = np.concatenate([recv_bg, recv_an])
    donors = np.concatenate([d1, d2])
    weights = np.concatenate([w1, w2])
    return m, recv, donors, weights
