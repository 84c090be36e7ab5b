"""SpherePack (lib/sphere-pack/SpherePack.hpp:20-120, py/pack.py) for building scenes:
makeCloud runs on the device (wdm_make_cloud), sphere after sphere in the reference's order,
so its result is the reference algorithm's on the same generator state.

The reference's generator is a function-level static (SpherePack.cpp:136-137): the first
makeCloud call of a process seeds it (seed, or the clock when seed is 0) and later calls go
on from where it stopped, whatever seed they pass.  This module keeps that state.
"""
import math
import time

import numpy as np

from . import soa

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
    """A solid that regularOrtho/regularHexa clip their lattice to (py/_packPredicates.cpp);
    `kind` and `params` are what the device evaluates."""
    kind = -1
    params = ()


class inAlignedBox(Predicate):
    """Axis-aligned box (py/_packPredicates.cpp:136-148)"""
    kind = 0

    def __init__(self, minAABB, maxAABB):
        self.params = tuple(float(x) for x in minAABB) + tuple(float(x) for x in maxAABB)

    def aabb(self):
        return self.params[:3], self.params[3:]


class inSphere(Predicate):
    """Sphere (py/_packPredicates.cpp:96-104)"""
    kind = 1

    def __init__(self, center, radius):
        self.params = tuple(float(x) for x in center) + (float(radius),)


class inCylinder(Predicate):
    """Cylinder between the centres of its bases (py/_packPredicates.cpp:189-205)"""
    kind = 2

    def __init__(self, centerBottom, centerTop, radius):
        self.params = tuple(float(x) for x in centerBottom) + tuple(float(x) for x in centerTop) + (float(radius),)


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
