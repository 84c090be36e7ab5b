"""SpherePack (lib/sphere-pack/SpherePack.hpp:20-120, py/pack.py) for building scenes:
makeCloud runs on the device (wdm_make_cloud), sphere after sphere in the reference's order,
so its result is the reference algorithm's on the same generator state.

The reference's generator is a function-level static (SpherePack.cpp:136-137): the first
makeCloud call of a process seeds it (seed, or the clock when seed is 0) and later calls go
on from where it stopped, whatever seed they pass.  This module keeps that state.
"""
import logging
import math
import time

import numpy as np

from . import soa

_log = logging.getLogger(__name__)
_M = 2147483647
_state = None


def _generator(seed):
    global _state
    if _state is None:
        s = (int(seed) if seed != 0 else int(time.time())) % _M
        _state = s if s != 0 else 1
    return np.array([_state], np.uint32)


def reset_generator():
    """forget the generator state (the next makeCloud seeds it again)"""
    global _state
    _state = None


class SpherePack:
    """A set of spheres (centres `c`, radii `r`); `cellSize` is nonzero for periodic packs."""

    def __init__(self):
        self.c = np.zeros((0, 3))
        self.r = np.zeros(0)
        self.clumpId = np.zeros(0, np.int64)  # -1: not in a clump
        self.cellSize = np.zeros(3)
        self.appliedPsdScaling = 1.
        self.userData = ""

    def __len__(self):
        return len(self.r)

    def __iter__(self):
        return iter(zip(map(tuple, self.c), self.r))

    def makeCloud(self, minCorner=(0., 0., 0.), maxCorner=(0., 0., 0.), rMean=-1, rRelFuzz=0, num=-1, periodic=False, porosity=.5,
                  psdSizes=(), psdCumm=(), distributeMass=False, seed=0, hSize=None, ctx=None, cap=None):
        """SpherePack::makeCloud (SpherePack.cpp:135-248, defaults of py/_packSpheres.cpp:29);
        returns the number of spheres.  With num < 0 the packing stops at `cap` spheres (by
        default the count that would fill the box densely) or at the first sphere without a
        free spot.  Sheared cells (hSize) are not supported."""
        if hSize is not None and np.any(np.asarray(hSize) != 0):
            raise NotImplementedError("makeCloud with hSize (sheared cells) is not supported")
        if len(self):
            raise NotImplementedError("makeCloud into a non-empty SpherePack is not supported")
        from ._native import Context
        global _state
        ctx = ctx or Context(0)
        mn, mx = np.asarray(minCorner, np.float64), np.asarray(maxCorner, np.float64)
        if cap is None:
            if num > 0:
                cap = num
            else:
                if len(psdSizes):
                    rmin = .5 * min(x for x in psdSizes if x > 0)
                else:
                    rm = rMean if rMean > 0 else 1.
                    rmin = rm * max(1e-3, 1 - abs(rRelFuzz))
                cap = int(min(2 ** 30, float(np.prod(mx - mn)) * .74 / (4 / 3. * math.pi * rmin ** 3) + 16))
        st = _generator(seed)
        c, r, sc = ctx.make_cloud(mn, mx, rMean, rRelFuzz, num, periodic, porosity, psdSizes, psdCumm, distributeMass, st, int(cap))
        _state = int(st[0])
        self.c, self.r, self.appliedPsdScaling = c, r, sc
        self.clumpId = np.full(len(r), -1, np.int64)
        self.cellSize = (mx - mn) if periodic else np.zeros(3)
        return len(r)

    def hasClumps(self):
        return bool(len(self.clumpId)) and bool(np.any(self.clumpId >= 0))

    def save(self, fileName):
        """SpherePack::toFile (SpherePack.cpp:114-129): optional `##PERIODIC:: x y z` and
        `##USER-DATA:: ...` lines, then `x y z r` (and the clump id when any sphere has one) per
        sphere, numbers as C++ streams write them by default (%g, 6 significant digits)."""
        if "\n" in self.userData:
            raise RuntimeError("SpherePack.userData must not contain newline.")
        clumps = self.hasClumps()
        with open(fileName, "w") as f:
            if np.any(self.cellSize != 0):
                f.write("##PERIODIC:: %g %g %g\n" % tuple(self.cellSize))
            if self.userData:
                f.write("##USER-DATA:: %s\n" % self.userData)
            for i in range(len(self.r)):
                f.write("%g %g %g %g" % (self.c[i, 0], self.c[i, 1], self.c[i, 2], self.r[i]))
                if clumps:
                    f.write(" %d" % self.clumpId[i])
                f.write("\n")

    saveTxt = save

    def load(self, fileName):
        """SpherePack::fromFile (SpherePack.cpp:73-112): the current spheres are replaced; every
        sphere line has 4 columns, or 5 with a clump id, all lines alike."""
        import os
        if not os.path.exists(fileName):
            raise ValueError("File with spheres `%s' doesn't exist." % fileName)
        c, r, cid, ncols = [], [], [], 0
        self.cellSize = np.zeros(3)
        with open(fileName) as f:
            for lineNo, line in enumerate(f, 1):
                line = line.rstrip("\n")
                if line.startswith("##USER-DATA:: "):
                    self.userData = line[14:]
                    continue
                tok = line.replace("\t", " ").split()
                if not tok:
                    continue
                if tok[0] == "##PERIODIC::":
                    if len(tok) != 4:
                        raise ValueError("Spheres file %s:%d contains ##PERIODIC::, but the line is malformed." % (fileName, lineNo))
                    self.cellSize = np.array([float(x) for x in tok[1:]])
                    continue
                if len(tok) not in (4, 5):
                    raise ValueError("%s:%d: line has %d columns (must be 4 or 5)." % (fileName, lineNo, len(tok)))
                ncols = ncols or len(tok)
                if len(tok) != ncols:
                    raise ValueError("%s:%d: all line must have the same number of columns (previous had %d, this one has %d)"
                                     % (fileName, lineNo, ncols, len(tok)))
                c.append([float(x) for x in tok[:3]])
                r.append(float(tok[3]))
                cid.append(int(tok[4]) if len(tok) == 5 else -1)
        self.c = np.array(c, np.float64).reshape(-1, 3)
        self.r = np.array(r, np.float64)
        self.clumpId = np.array(cid, np.int64)

    loadTxt = load

    def aabb(self):
        """SpherePack::aabb (SpherePack.hpp:106-110): the box around every sphere"""
        if not len(self.r):
            return np.full(3, np.inf), np.full(3, -np.inf)
        r = self.r[:, None]
        return (self.c - r).min(0), (self.c + r).max(0)

    def dim(self):
        mn, mx = self.aabb()
        return mx - mn

    def filtered(self, predicate, recenter=True):
        """SpherePack_filtered (py/_packPredicates.cpp:364-405): a new pack of the spheres for
        which predicate(c, r) holds (the radius is the pad); a clump goes whole when any of its
        spheres fails.  With recenter (and a finite predicate aabb) the spheres are first moved
        so that the centre of their aabb is the predicate's."""
        ret = SpherePack()
        ret.cellSize = np.array(self.cellSize, np.float64)
        off = np.zeros(3)
        smn, smx = self.aabb()
        pmn, pmx = (np.asarray(x, np.float64) for x in predicate.aabb())
        if recenter and not np.isinf((pmx - pmn).max()):
            off = (pmn + pmx) / 2 - (smn + smx) / 2
            smn, smx = smn + off, smx + off
        for ax in range(3):
            if pmn[ax] != -np.inf and pmn[ax] < smn[ax]:
                _log.warning("SpherePack.filtered: axis=%d, packing aabb (min=%g) outside of the predicate aabb (min=%g)", ax, smn[ax], pmn[ax])
            if pmx[ax] != np.inf and pmx[ax] > smx[ax]:
                _log.warning("SpherePack.filtered: axis=%d, packing aabb (max=%g) outside of the predicate aabb (max=%g)", ax, smx[ax], pmx[ax])
        c = self.c + off
        keep = np.asarray(predicate.contains(c, self.r), bool)
        cid = np.asarray(self.clumpId, np.int64) if len(self.clumpId) == len(self.r) else np.full(len(self.r), -1, np.int64)
        clumped = cid >= 0
        if np.any(clumped):
            gone = np.unique(cid[clumped & ~keep])
            keep = np.where(clumped, ~np.isin(cid, gone), keep)
        ret.c, ret.r, ret.clumpId = c[keep], np.array(self.r[keep]), cid[keep]
        return ret

    def toArrays(self, density=1e3, young=1e7, ktDivKn=.2, tanPhi=.5, mask=soa.DEFAULT_MOVABLE_MASK):
        """the spheres as scene arrays (free spheres, one FrictMat)"""
        return soa.spheres_arrays(self.c, self.r, density=density, young=young, ktDivKn=ktDivKn, tanPhi=tanPhi, mask=mask)

    def toSimulation(self, S, mat=None, mask=None):
        """add the spheres to S.dem (SpherePack.toSimulation, py/pack.py); returns their ids.
        Clumped spheres (clump ids) are not supported by this front end: they raise."""
        if self.hasClumps():
            raise NotImplementedError("toSimulation of a SpherePack with clumps")
        from . import utils
        from .dem import DemField
        kw = dict(mask=DemField.defaultMovableMask if mask is None else mask)
        if mat is not None:
            kw["mat"] = mat
        if np.any(self.cellSize != 0):
            S.periodic = True
            S.cell.setBox(self.cellSize)
        return S.dem.par.add([utils.sphere(tuple(c), float(r), **kw) for c, r in zip(self.c, self.r)])


# ---------------- predicates and regular packings ----------------
class Predicate:
    """A solid that packings are clipped to (py/_packPredicates.cpp); `kind` and `params` are
    what the device evaluates.  `pred(pt, pad)` is Predicate::operator() for one point,
    `pred.contains(c, pad)` the same for an (n, 3) array of points (pad scalar or per point),
    `pred.aabb()` the (min, max) corners."""
    kind = -1
    params = ()

    def __call__(self, pt, pad=0.):
        return bool(self.contains(np.asarray(pt, np.float64).reshape(1, 3), pad)[0])

    def center(self):
        mn, mx = self.aabb()
        return .5 * (np.asarray(mn) + np.asarray(mx))

    def dim(self):
        mn, mx = self.aabb()
        return np.asarray(mx) - np.asarray(mn)


class inAlignedBox(Predicate):
    """Axis-aligned box (py/_packPredicates.cpp:136-148)"""
    kind = 0

    def __init__(self, minAABB, maxAABB):
        self.params = tuple(float(x) for x in minAABB) + tuple(float(x) for x in maxAABB)

    def aabb(self):
        return self.params[:3], self.params[3:]

    def contains(self, c, pad=0.):
        p, c = self.params, np.asarray(c, np.float64).reshape(-1, 3)
        x, y, z = c[:, 0], c[:, 1], c[:, 2]
        return (p[0] + pad <= x) & (p[3] - pad >= x) & (p[1] + pad <= y) & (p[4] - pad >= y) & (p[2] + pad <= z) & (p[5] - pad >= z)


class inSphere(Predicate):
    """Sphere (py/_packPredicates.cpp:96-104).  The reference subtracts the pad on both sides
    of its test, (pt-center).norm()-pad <= radius-pad, so a point is inside when the centre is,
    whatever the pad; kept as it is."""
    kind = 1

    def __init__(self, center, radius):
        self.params = tuple(float(x) for x in center) + (float(radius),)

    def aabb(self):
        p = self.params
        return tuple(p[d] - p[3] for d in range(3)), tuple(p[d] + p[3] for d in range(3))

    def contains(self, c, pad=0.):
        p, c = self.params, np.asarray(c, np.float64).reshape(-1, 3)
        dx, dy, dz = c[:, 0] - p[0], c[:, 1] - p[1], c[:, 2] - p[2]
        return np.sqrt((dx * dx + dy * dy) + dz * dz) - pad <= p[3] - pad


class inCylinder(Predicate):
    """Cylinder between the centres of its bases (py/_packPredicates.cpp:189-213)"""
    kind = 2

    def __init__(self, centerBottom, centerTop, radius):
        self.params = tuple(float(x) for x in centerBottom) + tuple(float(x) for x in centerTop) + (float(radius),)
        c12 = np.subtract(centerTop, centerBottom, dtype=np.float64)
        if not math.sqrt((c12[0] * c12[0] + c12[1] * c12[1]) + c12[2] * c12[2]) > 0:
            raise ValueError("inCylinder: the axis has zero length")

    def _ht(self):
        p = self.params
        cx, cy, cz = p[3] - p[0], p[4] - p[1], p[5] - p[2]
        return cx, cy, cz, math.sqrt((cx * cx + cy * cy) + cz * cz)

    def aabb(self):
        # inCylinder::aabb: the bases' extreme points, radius * sin of the axis' angle per axis
        A, B, r = self.params[:3], self.params[3:6], self.params[6]
        ht = self._ht()[3]
        k = (math.sqrt((A[1] - B[1]) ** 2 + (A[2] - B[2]) ** 2) / ht, math.sqrt((A[0] - B[0]) ** 2 + (A[2] - B[2]) ** 2) / ht,
             math.sqrt((A[0] - B[0]) ** 2 + (A[1] - B[1]) ** 2) / ht)
        return tuple(min(A[d], B[d]) - r * k[d] for d in range(3)), tuple(max(A[d], B[d]) + r * k[d] for d in range(3))

    def contains(self, c, pad=0.):
        p, c = self.params, np.asarray(c, np.float64).reshape(-1, 3)
        cx, cy, cz, ht = self._ht()
        x, y, z = c[:, 0], c[:, 1], c[:, 2]
        u = (((x * cx + y * cy) + z * cz) - ((p[0] * cx + p[1] * cy) + p[2] * cz)) / (ht * ht)
        axial = ~((u * ht < 0 + pad) | (u * ht > ht - pad))
        ax, ay, az, bx, by, bz = x - p[0], y - p[1], z - p[2], x - p[3], y - p[4], z - p[5]
        qx, qy, qz = ay * bz - az * by, az * bx - ax * bz, ax * by - ay * bx
        axisDist = np.sqrt((qx * qx + qy * qy) + qz * qz) / ht
        return axial & ~(axisDist > p[6] - pad)


def _regular(hexa, predicate, radius, gap, ctx, cap, kw):
    from . import utils
    from ._native import Context
    if not isinstance(predicate, Predicate) or predicate.kind < 0:
        raise NotImplementedError("regular packings support the predicates inAlignedBox, inSphere and inCylinder")
    ctx = ctx or Context(0)
    c = ctx.regular_pack(hexa, predicate.kind, predicate.params, radius, gap, cap)
    return [utils.sphere(tuple(x), radius=radius, **kw) for x in c]


def regularOrtho(predicate, radius, gap, ctx=None, cap=None, **kw):
    """Spheres on an orthogonal grid clipped to the predicate, separated by `gap`
    (py/pack.py:323-332), computed on the device; kw go to utils.sphere."""
    return _regular(False, predicate, radius, gap, ctx, cap, kw)


def regularHexa(predicate, radius, gap, ctx=None, cap=None, **kw):
    """Spheres on a hexagonal grid clipped to the predicate, separated by `gap`
    (py/pack.py:334-350), computed on the device; kw go to utils.sphere."""
    return _regular(True, predicate, radius, gap, ctx, cap, kw)


def filterSpherePack(predicate, spherePack, **kw):
    """py/pack.py:387-391: spherePack.filtered(predicate)"""
    return spherePack.filtered(predicate)


def sanitizePsd(psdPts, src="psdPts"):
    """PsdSphereGenerator::sanitizePsd (pkg/dem/Psd.cpp:19-40): (diameter, passing) points,
    both non-decreasing and non-negative; leading zero-passing points but the last of them and
    trailing points at the top passing value but the first of them dropped; passing scaled so
    that the top is 1."""
    pts = [(float(d), float(p)) for d, p in psdPts]
    if not pts:
        return pts
    for i in range(len(pts) - 1):
        if pts[i][0] < 0 or pts[i][1] < 0:
            raise RuntimeError(src + ": negative values not allowed.")
        if pts[i][0] > pts[i + 1][0]:
            raise RuntimeError(src + ": diameters (the x-component) must be increasing (%g>=%g)" % (pts[i][0], pts[i + 1][0]))
        if pts[i][1] > pts[i + 1][1]:
            raise RuntimeError(src + ": passing values (the y-component) must be increasing (%g>%g)" % (pts[i][1], pts[i + 1][1]))
    maxPass = pts[-1][1]
    ii = 0
    while ii < len(pts) and not pts[ii][1] > 0:
        ii += 1
    if ii > 1:
        pts = pts[ii - 1:]
    ii = len(pts) - 2
    while ii >= 0 and not pts[ii][1] < maxPass:
        ii -= 1
    if ii < len(pts) - 2:
        pts = pts[:ii + 2]
    if not pts:
        raise RuntimeError(src + ": empty after removing spurious values?")
    if maxPass != 1.:
        pts = [(d, p / maxPass) for d, p in pts]
    return pts


def randomLoosePsd(predicate, psd, mass=True, discrete=False, maxAttempts=5000, clumps=[], returnSpherePack=False, seed=0, ctx=None,
                   cap=None, **kw):
    """A loose random packing with the particle size distribution `psd` ((diameter, passing)
    points) filling the predicate (py/pack.py:353-380), as a SpherePack.

    The reference fills the predicate's aabb with a BoxInlet of a PsdSphereGenerator (always
    by mass and continuous, whatever `mass` and `discrete` say) until `maxAttempts` placements
    in a row fail, then keeps the spheres inside the predicate (SpherePack.filtered).  Here the
    aabb is filled on the device by makeCloud with the sanitized PSD distributed by mass (the
    random sequential placement of SpherePack.cpp:135-248, which gives up on a sphere after its
    own 1000 attempts and stops there), then filtered alike.  Both are random loose packings of
    the same PSD in the same region; the generators differ, so the spheres are not the same.
    As in the reference, returnSpherePack must be True; clumps are not supported."""
    if clumps:
        raise NotImplementedError("randomLoosePsd with clumps is not supported")
    if not returnSpherePack:
        raise ValueError("returnSpherePack must be True.")
    pts = sanitizePsd(psd, "PsdSphereGenerator.psdPts")
    if not pts:
        raise RuntimeError("PsdSphereGenerator.psdPts is empty.")
    mn, mx = predicate.aabb()
    if not (np.all(np.isfinite(mn)) and np.all(np.isfinite(mx))):
        raise ValueError("randomLoosePsd: the predicate's aabb must be finite")
    sizes, cumm = [p[0] for p in pts], [p[1] for p in pts]
    if cumm[0] != 0.:
        # the generator makes the passing of the first point spheres of exactly that diameter
        # (Psd.cpp:96); makeCloud gets it as a bin one millionth of the diameter wide
        sizes, cumm = [sizes[0] * (1 - 1e-6)] + sizes, [0.] + cumm
    sp = SpherePack()
    sp.makeCloud(mn, mx, psdSizes=sizes, psdCumm=cumm, distributeMass=True, seed=seed, ctx=ctx, cap=cap)
    return sp.filtered(predicate)
