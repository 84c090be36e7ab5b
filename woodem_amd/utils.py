"""Particle factories and default engines (module woo.utils of the reference, py/utils.py)."""
import math
import random

import numpy as np

from . import soa as _soa
from .dem import (Bo1_Facet_Aabb, Bo1_Sphere_Aabb, Bo1_Wall_Aabb, Cg2_Facet_Sphere_L6Geom, Cg2_Sphere_Sphere_L6Geom,
                  Cg2_Wall_Sphere_L6Geom, ContactLoop, Cp2_FrictMat_FrictPhys, DemData, DemField, DynDt, Facet, FrictMat,
                  InsertionSortCollider, Law2_L6Geom_FrictPhys_IdealElPl, Leapfrog, Node, Particle, Sphere, Wall)

inf = float("inf")


def spherePWaveDt(radius, density, young):
    """P-wave critical timestep r/sqrt(E/rho), py/utils.py:38-46."""
    return radius / math.sqrt(young / density)


def defaultMaterial():
    """FrictMat(density=1e3, young=1e7, ktDivKn=.2, tanPhi=tan(.5)), py/utils.py:48-53."""
    return FrictMat(density=1e3, young=1e7, ktDivKn=.2, tanPhi=math.tan(.5))


def defaultEngines(damping=0., gravity=None, verletDist=-.05, kinSplit=False, dontCollect=False, noSlip=False, noBreak=False,
                   cp2=None, law=None, model=None, grid=False, dynDtPeriod=100, cpKw={}, lawKw={}, precision="fp64"):
    """Canonical timestep engines, py/utils.py:55-83: Leapfrog(reset=True), InsertionSortCollider,
    ContactLoop, and DynDt(stepPeriod=dynDtPeriod) when dynDtPeriod > 0."""
    if gravity:
        raise ValueError("gravity MUST NOT be specified anymore, set DemField.gravity=... instead.")
    if model or grid:
        raise NotImplementedError("model= and grid= are outside this path")
    cp2 = [cp2 if cp2 else Cp2_FrictMat_FrictPhys()]
    law = [law if law else Law2_L6Geom_FrictPhys_IdealElPl(noSlip=noSlip, noBreak=noBreak, label="contactLaw")]
    cp2[0].updateAttrs(cpKw)
    law[0].updateAttrs(lawKw)
    distFactor = 1.
    collider = InsertionSortCollider([Bo1_Sphere_Aabb(distFactor=distFactor), Bo1_Facet_Aabb(), Bo1_Wall_Aabb()],
                                     label="collider", verletDist=verletDist)
    return [
        Leapfrog(damping=damping, reset=True, kinSplit=kinSplit, dontCollect=dontCollect, label="leapfrog"),
        collider,
        ContactLoop([Cg2_Sphere_Sphere_L6Geom(distFactor=distFactor), Cg2_Facet_Sphere_L6Geom(), Cg2_Wall_Sphere_L6Geom()],
                    cp2, law, applyForces=True, label="contactLoop", precision=precision),
    ] + ([DynDt(stepPeriod=dynDtPeriod, label="dynDt")] if dynDtPeriod > 0 else [])


def _mat(mat):
    if callable(mat) and not isinstance(mat, FrictMat):
        return mat()
    return mat


def _node(pos):
    return Node(pos=np.asarray(pos, dtype=np.float64).copy(), dem=DemData())


def sphere(center, radius, mat=defaultMaterial, fixed=False, wire=False, color=None, highlight=False,
           mask=DemField.defaultMovableMask):
    """Sphere with mass/inertia from Sphere::lumpMassInertia (py/utils.py:107-152, pkg/dem/Sphere.cpp:24-31)."""
    m = _mat(mat)
    p = Particle(shape=Sphere(radius=radius, color=color if color else random.random()), material=m, mask=mask)
    n = _node(center)
    mass, I = _soa.sphere_mass_inertia(radius, m.density)
    n.dem.mass = float(mass)
    n.dem.inertia = np.full(3, float(I))
    if fixed:
        n.dem.blocked = "xyzXYZ"
    p.nodes = [n]
    return p


def wall(position, axis, sense=0, glAB=None, fixed=True, mass=0, color=None, mat=defaultMaterial, visible=True,
         mask=DemField.defaultBoundaryMask):
    """Axis-aligned wall, py/utils.py:177-199."""
    if not fixed and mass <= 0:
        raise ValueError("Non-fixed wall must have positive mass")
    if isinstance(position, (int, float)):
        pos = np.zeros(3)
        pos[axis] = position
    else:
        pos = np.asarray(position, dtype=np.float64)
    p = Particle(shape=Wall(sense=sense, axis=axis, color=color if color else random.random(), glAB=glAB), material=_mat(mat),
                 mask=mask)
    n = _node(pos)
    n.dem.mass = mass
    if fixed:
        n.dem.blocked = "xyzXYZ"
    p.nodes = [n]
    return p


def wallBox(box, which=(1, 1, 1, 1, 1, 1), **kw):
    """Walls -x,+x,-y,+y,-z,+z of an axis-aligned box (py/utils.py:201-210): sense +1 on the
    minimum faces, -1 on the maximum faces."""
    mn, mx = np.asarray(box[0], dtype=np.float64), np.asarray(box[1], dtype=np.float64)
    ret = []
    for sense, ix, coord in [(1, 0, mn), (-1, 1, mx)]:
        for ax in (0, 1, 2):
            if not which[3 * ix + ax]:
                continue
            ret.append(wall(float(coord[ax]), axis=ax, sense=sense, **kw))
    return ret


def facet(vertices, fakeVel=None, halfThick=0., fixed=True, wire=True, color=None, highlight=False, mat=defaultMaterial,
          visible=True, mask=DemField.defaultBoundaryMask):
    """Triangular facet, py/utils.py:242-268; vertices may be Node objects shared between facets."""
    if fakeVel is not None:
        raise NotImplementedError("Facet.fakeVel is outside this path")
    nodes = [v if isinstance(v, Node) else _node(v) for v in vertices]
    for n in nodes:
        if n.dem is None:
            n.dem = DemData()
        if fixed:
            n.dem.blocked = "xyzXYZ"
    p = Particle(shape=Facet(halfThick=halfThick, color=color if color else random.random()), material=_mat(mat), mask=mask)
    p.nodes = nodes
    return p
