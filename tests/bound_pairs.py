"""Potential contacts of a periodic sphere packing right after a collider full run, computed
without the collider: every pair of spheres whose bounds overlap in the periodic cell
(InsertionSortCollider::spatialOverlapPeri, pkg/dem/InsertionSortCollider.cpp:903-957, with
cellWrap / cellWrapRel of :650-660), the bounds being Bo1_Sphere_Aabb's (pkg/dem/Sphere.cpp:50-79:
centre -+ distFactor r) enlarged by the Verlet distance at the full run
(InsertionSortCollider::updateBboxes_doFullRun, :374-398).  At a full run every bound is
recomputed from the current positions, so the positions and the cell read back right after
it fix the set.

Candidates come from periodic k-d trees (scipy cKDTree, Chebyshev distance), one query per
pair of radius classes, so that the few large spheres do not widen every query; the exact
predicate then runs in float64 with the collider's own operations in their order, so the set
is the one the collider must produce, at any size (1,000,000 spheres in seconds)."""
import numpy as np
from scipy.spatial import cKDTree


def sphere_bounds(pos, radius, distFactor, verletDist):
    """Bo1_Sphere_Aabb, then the Verlet enlargement: (x - h) - vd and (x + h) + vd."""
    h = (distFactor if distFactor > 0 else 1.) * np.asarray(radius, np.float64)
    mn = pos - h[:, None]
    mx = pos + h[:, None]
    if verletDist > 0:
        mn = mn - verletDist
        mx = mx + verletDist
    return mn, mx


def _wrap(x, x0, x1):
    xn = (x - x0) / (x1 - x0)
    p = np.floor(xn)
    return x0 + (xn - p) * (x1 - x0), p


def _wrap_rel(x, x0, x1):
    xn = (x - x0) / (x1 - x0)
    return (xn - np.floor(xn)) * (x1 - x0)


def overlap_peri(mn, mx, i, j, size):
    """spatialOverlapPeri(i, j) for arrays of pairs: (overlap, periods (n, 3))."""
    ok = np.ones(len(i), bool)
    per = np.zeros((len(i), 3), np.int64)
    for ax in range(3):
        dim = float(size[ax])
        m1, m2 = mn[i, ax], mn[j, ax]
        wmn = np.where(_wrap_rel(m1, m2, m2 + dim) < _wrap_rel(m2, m1, m1 + dim), m2, m1)
        lo1, p1 = _wrap(m1, wmn, wmn + dim)
        hi1, q1 = _wrap(mx[i, ax], wmn, wmn + dim)
        lo2, p2 = _wrap(m2, wmn, wmn + dim)
        hi2, q2 = _wrap(mx[j, ax], wmn, wmn + dim)
        if not (np.array_equal(p1, q1) and np.array_equal(p2, q2)):
            raise ValueError("a bound longer than the cell")
        per[:, ax] = (p1 - p2).astype(np.int64)
        ok &= (lo1 <= hi2) & (hi1 >= lo2)
    return ok, per


def _candidates(pos, half, size, classes=4):
    """(i, j), i < j: pairs whose periodic centre distance along every axis is at most the sum
    of the two half-widths (with a margin), a superset of the overlapping pairs."""
    n = len(pos)
    w = np.mod(pos, size[None, :])
    w = np.where(w >= size[None, :], 0., w)
    edges = np.quantile(half, np.linspace(0, 1, classes + 1))
    cls = np.clip(np.searchsorted(edges[1:-1], half, side="right"), 0, classes - 1)
    members = [np.flatnonzero(cls == c) for c in range(classes)]
    trees = [cKDTree(w[m], boxsize=size) if len(m) else None for m in members]
    hmax = [float(half[m].max()) if len(m) else 0. for m in members]
    out_i, out_j = [], []
    for a in range(classes):
        for b in range(a, classes):
            if trees[a] is None or trees[b] is None:
                continue
            r = (hmax[a] + hmax[b]) * (1 + 1e-9) + 1e-15
            if a == b:
                pr = trees[a].query_pairs(r, p=np.inf, output_type="ndarray")
                ia, ib = members[a][pr[:, 0]], members[a][pr[:, 1]]
            else:
                sd = trees[a].sparse_distance_matrix(trees[b], r, p=np.inf, output_type="ndarray")
                ia, ib = members[a][sd["i"]], members[b][sd["j"]]
            out_i.append(np.minimum(ia, ib))
            out_j.append(np.maximum(ia, ib))
    if not out_i:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    i, j = np.concatenate(out_i).astype(np.int64), np.concatenate(out_j).astype(np.int64)
    assert n < (1 << 31)
    return i, j


def periodic_bound_pairs(pos, radius, distFactor, verletDist, size):
    """Sorted int64 keys (min id << 32 | max id) of the sphere pairs whose enlarged bounds
    overlap in a periodic cell with edge lengths `size` (non-sheared)."""
    size = np.asarray(size, np.float64)
    mn, mx = sphere_bounds(pos, radius, distFactor, verletDist)
    half = (mx - mn).max(axis=1) * .5
    i, j = _candidates(pos, half, size)
    ok, _ = overlap_peri(mn, mx, i, j, size)
    return np.sort((i[ok] << 32) | j[ok])


def contact_periods(pos, radius, distFactor, verletDist, size, idA, idB):
    """(overlap, periods) of spatialOverlapPeri(idA, idB) for given contacts."""
    mn, mx = sphere_bounds(pos, radius, distFactor, verletDist)
    return overlap_peri(mn, mx, np.asarray(idA, np.int64), np.asarray(idB, np.int64), np.asarray(size, np.float64))


def cell_size(hSize):
    """Cell edge lengths: the norms of hSize's columns (Cell::integrateAndUpdate, core/Cell.cpp:101)."""
    h = np.asarray(hSize, np.float64).reshape(3, 3)
    return np.sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2])
