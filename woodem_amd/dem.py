"""Woo's dem plugin vocabulary for the timestep path (module woo.dem of the reference).

Classes keep the reference's names and attributes; they only describe the
scene and its engines.  The arithmetic runs in libwoodem_hip.so, driven by
woodem_amd.core.Scene.

References: Particle/DemData/DemField pkg/dem/Particle.hpp:26-459; shapes
pkg/dem/Sphere.hpp, Wall.hpp, Facet.hpp; FrictMat pkg/dem/FrictMat.hpp:18-28;
engines pkg/dem/Leapfrog.hpp:42-61, InsertionSortCollider.hpp:252-265,
ContactLoop.hpp:108-128; functors pkg/dem/Sphere.hpp, Wall.hpp, Facet.hpp,
L6Geom.hpp:29-44, FrictMat.hpp, IdealElPl.hpp:10-26, Hertz.hpp:27-72.
"""
import math

import numpy as np

from . import soa as _soa
from ._native import SHAPE_FACET, SHAPE_SPHERE, SHAPE_WALL


class _Attrs:
    """Keyword-attribute object, like Woo's Object(**kw) constructors."""

    _defaults = {}

    def __init__(self, **kw):
        for k, v in self._defaults.items():
            setattr(self, k, v() if callable(v) else v)
        for k, v in kw.items():
            if k not in self._defaults and k != "label":
                raise AttributeError("%s has no attribute %r" % (type(self).__name__, k))
            setattr(self, k, v)
        if not hasattr(self, "label"):
            self.label = ""

    def updateAttrs(self, d):
        for k, v in d.items():
            setattr(self, k, v)

    def __repr__(self):
        return "<%s>" % type(self).__name__


# ---------------- materials ----------------
class ElastMat(_Attrs):
    _defaults = dict(density=1e3, young=1e9)


class FrictMat(ElastMat):
    """FrictMat, pkg/dem/FrictMat.hpp:18-28 (tanPhi=.5, ktDivKn=.2)."""
    _defaults = dict(density=1e3, young=1e9, tanPhi=.5, ktDivKn=.2)


# ---------------- shapes ----------------
class Shape(_Attrs):
    nNodes = 1


class Sphere(Shape):
    _defaults = dict(radius=float("nan"), color=0.5, wire=False)


class Wall(Shape):
    """Axis-aligned infinite plane, pkg/dem/Wall.hpp:18-20: sense -1/0/+1, axis 0/1/2."""
    _defaults = dict(sense=0, axis=0, color=0.5, glAB=None)


class Facet(Shape):
    nNodes = 3
    _defaults = dict(halfThick=0., color=0.5, wire=True, fakeVel=None)


# ---------------- nodes and particles ----------------
class DemData(_Attrs):
    """Dynamic state of a node, pkg/dem/Particle.hpp:249-262."""
    _defaults = dict(vel=lambda: np.zeros(3), angVel=lambda: np.zeros(3), mass=0., inertia=lambda: np.zeros(3),
                     force=lambda: np.zeros(3), torque=lambda: np.zeros(3), flags=0)

    @property
    def blocked(self):
        return "".join(c for i, c in enumerate("xyzXYZ") if self.flags & (1 << i))

    @blocked.setter
    def blocked(self, s):
        self.flags &= ~_soa.DOF_ALL
        for c in s:
            self.flags |= 1 << "xyzXYZ".index(c)

    gravitySkip = property(lambda s: bool(s.flags & _soa.GRAVITY_SKIP),
                           lambda s, v: setattr(s, "flags", (s.flags | _soa.GRAVITY_SKIP) if v else (s.flags & ~_soa.GRAVITY_SKIP)))
    dampingSkip = property(lambda s: bool(s.flags & _soa.DAMPING_SKIP),
                           lambda s, v: setattr(s, "flags", (s.flags | _soa.DAMPING_SKIP) if v else (s.flags & ~_soa.DAMPING_SKIP)))
    energySkip = property(lambda s: bool(s.flags & _soa.ENERGY_SKIP),
                          lambda s, v: setattr(s, "flags", (s.flags | _soa.ENERGY_SKIP) if v else (s.flags & ~_soa.ENERGY_SKIP)))


class Node(_Attrs):
    _defaults = dict(pos=lambda: np.zeros(3), ori=lambda: np.array([1., 0, 0, 0]), dem=None)


class Particle(_Attrs):
    """Particle, pkg/dem/Particle.hpp:26-130."""
    _defaults = dict(shape=None, material=None, mask=_soa.DEFAULT_MOVABLE_MASK, nodes=lambda: [])

    def __init__(self, **kw):
        super().__init__(**kw)
        self.id = -1
        self._field = None

    mat = property(lambda s: s.material, lambda s, v: setattr(s, "material", v))

    def _node_ix(self):
        return self._field._node_index_of(self)

    def _get(self, what):
        f = self._field
        if f is None or f._arrays is None:
            n = self.nodes[0]
            return n.pos if what == "pos" else getattr(n.dem, what)
        f._refresh()
        return f._arrays_cur[what][self._node_ix()].copy()

    pos = property(lambda s: s._get("pos"))
    vel = property(lambda s: s._get("vel"))
    angVel = property(lambda s: s._get("angVel"))
    f = property(lambda s: s._get("force"))
    t = property(lambda s: s._get("torque"))

    @property
    def mass(self):
        return self.nodes[0].dem.mass

    @property
    def Ek(self):
        """Kinetic energy of the particle's node (translation + rotation)."""
        v = self.vel
        w = self.angVel
        d = self.nodes[0].dem
        return .5 * d.mass * float(np.dot(v, v)) + .5 * float(np.dot(w, np.asarray(d.inertia) * w))


class ParticleContainer:
    """DemField.par: list of particles with sequential ids (pkg/dem/ParticleContainer.hpp)."""

    def __init__(self, field):
        self._f = field
        self._p = []

    def add(self, p):
        if isinstance(p, (list, tuple)):
            return [self.add(x) for x in p]
        p.id = len(self._p)
        p._field = self._f
        self._p.append(p)
        self._f._dirty = True
        return p.id

    def __getitem__(self, i):
        return self._p[i]

    def __len__(self):
        return len(self._p)

    def __iter__(self):
        return iter(self._p)

    def exists(self, i):
        return 0 <= i < len(self._p) and self._p[i] is not None

    def _removed(self, ids):
        """particles removed on the device (DemField::removeParticle): their slots become None,
        as ParticleContainer::remove leaves them (ParticleContainer.cpp:121-126)"""
        for i in ids:
            self._p[int(i)] = None


class ContactView:
    """One contact of DemField.con (Contact + L6Geom + FrictPhys data, read-only)."""

    def __init__(self, d, i):
        self.id1, self.id2 = int(d["idA"][i]), int(d["idB"][i])
        self.real = bool(d["isReal"][i])
        self.cellDist = d["cellDist"][i].copy()
        self.force = d["forceLocal"][i].copy()
        self.contPt = d["contPt"][i].copy()
        self.trsf = d["trsf"][i].copy()
        self.uN = float(d["uN"][i])

    @property
    def ids(self):
        return (min(self.id1, self.id2), max(self.id1, self.id2))


class ContactContainer:
    """DemField.con: contacts (real and potential) after the last step."""

    def __init__(self, field):
        self._f = field

    def _data(self):
        return self._f._contacts()

    def __len__(self):
        return len(self._data()["idA"])

    def __getitem__(self, i):
        return ContactView(self._data(), i)

    def countReal(self):
        return int(np.count_nonzero(self._data()["isReal"]))

    def arrays(self):
        return self._data()

    def __iter__(self):  # real contacts only, as ContactContainer::pyIterator (ContactContainer.cpp:223-229)
        d = self._data()
        for i in np.nonzero(d["isReal"])[0]:
            yield ContactView(d, i)


class DemField(_Attrs):
    """DemField, pkg/dem/Particle.hpp:278-330."""
    _defaults = dict(gravity=lambda: np.zeros(3), loneMask=_soa.DEFAULT_LONE_MASK)
    defaultMovableMask = _soa.DEFAULT_MOVABLE_MASK
    defaultBoundaryMask = _soa.DEFAULT_BOUNDARY_MASK
    defaultLoneMask = _soa.DEFAULT_LONE_MASK

    def __init__(self, **kw):
        super().__init__(**kw)
        self.par = ParticleContainer(self)
        self.con = ContactContainer(self)
        self._arrays = None      # DemArrays given directly or compiled from par
        self._arrays_cur = None  # latest downloaded node state
        self._dirty = True
        self._scene = None
        self._node_of = None

    # bulk construction, bypassing per-particle objects (large generated scenes)
    def setArrays(self, arrays):
        self._arrays = arrays.normalized()
        self._arrays.validate()
        self._dirty = True
        self._from_arrays = True

    def _node_index_of(self, p):
        return self._node_of[p.id]

    def _compile(self):
        if getattr(self, "_from_arrays", False) and not len(self.par):
            return self._arrays
        slots = list(self.par)
        n_par = len(slots)
        pars = [p for p in slots if p is not None]  # holes (removed particles) become inert slots
        mats, matIx = [], {}
        for p in pars:
            if id(p.material) not in matIx:
                matIx[id(p.material)] = len(mats)
                mats.append(p.material)
        # node numbering: uninodal particles own node id == particle id; multinodal nodes follow
        node_ix = {}
        extra = []
        for p in pars:
            if isinstance(p.shape, (Sphere, Wall)):
                if len(p.nodes) != 1:
                    raise ValueError("particle %d must have one node" % p.id)
                if id(p.nodes[0]) in node_ix:
                    raise ValueError("uninodal particles may not share nodes on this path")
                node_ix[id(p.nodes[0])] = p.id
        nxt = n_par
        for p in pars:
            if isinstance(p.shape, Facet):
                for n in p.nodes:
                    if id(n) not in node_ix:
                        node_ix[id(n)] = nxt
                        extra.append(n)
                        nxt += 1
            elif not isinstance(p.shape, (Sphere, Wall)):
                raise ValueError("unsupported shape %s" % type(p.shape).__name__)
        nn = nxt
        allnodes = [None] * nn
        for p in pars:
            for n in p.nodes:
                allnodes[node_ix[id(n)]] = n
        for i, p in enumerate(slots):
            if p is None:  # the hole's node: fully blocked, massless, skipped by the energy terms
                allnodes[i] = Node(dem=DemData(flags=_soa.DOF_ALL | _soa.ENERGY_SKIP | _soa.GRAVITY_SKIP | _soa.DAMPING_SKIP))
        pos = np.array([n.pos for n in allnodes], dtype=np.float64).reshape(nn, 3)
        ori = np.array([n.ori for n in allnodes], dtype=np.float64).reshape(nn, 4)
        vel = np.array([n.dem.vel for n in allnodes], dtype=np.float64).reshape(nn, 3)
        angVel = np.array([n.dem.angVel for n in allnodes], dtype=np.float64).reshape(nn, 3)
        mass = np.array([n.dem.mass for n in allnodes], dtype=np.float64)
        inertia = np.array([n.dem.inertia for n in allnodes], dtype=np.float64).reshape(nn, 3)
        flags = np.array([n.dem.flags for n in allnodes], dtype=np.uint32)
        force = np.array([n.dem.force for n in allnodes], dtype=np.float64).reshape(nn, 3)
        torque = np.array([n.dem.torque for n in allnodes], dtype=np.float64).reshape(nn, 3)
        shape = np.zeros(n_par, np.int32)
        nodes = -np.ones((n_par, 3), np.int32)
        radius = np.zeros(n_par)
        axis = np.zeros(n_par, np.int32)
        sense = np.zeros(n_par, np.int32)
        mask = np.zeros(n_par, np.uint32)
        mat = np.zeros(n_par, np.int32)
        for p in pars:
            i = p.id
            s = p.shape
            if isinstance(s, Sphere):
                shape[i] = SHAPE_SPHERE
                radius[i] = s.radius
            elif isinstance(s, Wall):
                shape[i] = SHAPE_WALL
                axis[i], sense[i] = s.axis, s.sense
            else:
                shape[i] = SHAPE_FACET
                radius[i] = s.halfThick
            for k, n in enumerate(p.nodes):
                nodes[i, k] = node_ix[id(n)]
            mask[i] = p.mask
            mat[i] = matIx[id(p.material)]
        self._node_of = np.array([node_ix[id(p.nodes[0])] if p is not None else i for i, p in enumerate(slots)], dtype=np.int64)
        arr = _soa.DemArrays(pos=pos, ori=ori, vel=vel, angVel=angVel, mass=mass, inertia=inertia, flags=flags,
                             force=force, torque=torque, shape=shape, nodes=nodes, radius=radius, axis=axis, sense=sense,
                             mask=mask, mat=mat,
                             density=np.array([m.density for m in mats], dtype=np.float64),
                             young=np.array([m.young for m in mats], dtype=np.float64),
                             ktDivKn=np.array([getattr(m, "ktDivKn", .2) for m in mats], dtype=np.float64),
                             tanPhi=np.array([getattr(m, "tanPhi", .5) for m in mats], dtype=np.float64)).normalized()
        arr.validate()
        self._arrays = arr
        return arr

    def _mark_removed(self, ids):
        """the compiled arrays follow a removal on the device (shape NONE, inert node)"""
        a = self._arrays
        a.shape[ids] = 0
        a.flags[a.nodes[ids, 0]] = _soa.DOF_ALL | _soa.ENERGY_SKIP | _soa.GRAVITY_SKIP | _soa.DAMPING_SKIP
        a.mass[a.nodes[ids, 0]] = 0.

    def _refresh(self):
        if self._scene is not None:
            self._scene._download_if_stale()

    def _contacts(self):
        if self._scene is None:
            raise RuntimeError("no contacts before the scene has run")
        return self._scene._contacts()

    def nodeArrays(self):
        """Current node state as numpy arrays (pos, vel, angVel, ori, force, torque)."""
        self._refresh()
        return self._arrays_cur

    def critDt(self):
        """DemField::critDt = DemFuncs::pWaveDt (pkg/dem/Funcs.cpp:116-142, Particle.cpp:372-377)."""
        a = self._arrays if (self._arrays is not None and not self._dirty) else self._compile()
        sph = a.shape == SHAPE_SPHERE
        if not np.any(sph):
            return float("inf")
        r = a.radius[sph]
        m = a.mat[sph]
        return float(np.min(r / np.sqrt(a.young[m] / a.density[m])))


# ---------------- functors ----------------
class Bo1_Sphere_Aabb(_Attrs):
    _defaults = dict(distFactor=1.)


class Bo1_Wall_Aabb(_Attrs):
    pass


class Bo1_Facet_Aabb(_Attrs):
    pass


class Cg2_Any_Any_L6Geom__Base(_Attrs):
    """pkg/dem/L6Geom.hpp:29-44 (noRatch, approxMask must stay at their defaults on this path)."""
    _defaults = dict(noRatch=False, trsfRenorm=100, approxMask=0)


class Cg2_Sphere_Sphere_L6Geom(Cg2_Any_Any_L6Geom__Base):
    _defaults = dict(Cg2_Any_Any_L6Geom__Base._defaults, distFactor=1.)


class Cg2_Wall_Sphere_L6Geom(Cg2_Any_Any_L6Geom__Base):
    pass


class Cg2_Facet_Sphere_L6Geom(Cg2_Any_Any_L6Geom__Base):
    pass


class Cp2_FrictMat_FrictPhys(_Attrs):
    _defaults = dict(tanPhi=None)


class Cp2_FrictMat_HertzPhys(Cp2_FrictMat_FrictPhys):
    """pkg/dem/Hertz.hpp:27-32."""
    _defaults = dict(tanPhi=None, poisson=.2, gamma=0., en=float("nan"), alpha=0.)


class Law2_L6Geom_FrictPhys_IdealElPl(_Attrs):
    """pkg/dem/IdealElPl.hpp:10-26."""
    _defaults = dict(iniEqlb=False, relRollStiff=0., relTwistStiff=0., rollTanPhi=0., noSlip=False, noBreak=False,
                     noFrict=False)


class Law2_L6Geom_HertzPhys_DMT(_Attrs):
    _defaults = dict(noAttraction=True)


# ---------------- engines ----------------
class Engine(_Attrs):
    dead = False


class Leapfrog(Engine):
    """pkg/dem/Leapfrog.hpp:42-61."""
    _defaults = dict(damping=.2, reset=False, kinSplit=False, dontCollect=False)


class ForceResetter(Engine):
    pass


class InsertionSortCollider(Engine):
    """pkg/dem/InsertionSortCollider.hpp:252-265; the positional argument is the list of Bo1 functors."""
    _defaults = dict(verletDist=-.05, sortAxis=0, forceInitSort=False, sortThenCollide=False, noBoundOk=False,
                     boundFunctors=lambda: [])

    def __init__(self, boundFunctors=None, **kw):
        super().__init__(**kw)
        if boundFunctors is not None:
            self.boundFunctors = list(boundFunctors)


class ContactLoop(Engine):
    """pkg/dem/ContactLoop.hpp:108-128; positional arguments: Cg2, Cp2, Law2 functor lists.
    `precision` ('fp64' or 'mixed') selects the arithmetic of the contact kernel."""
    _defaults = dict(applyForces=True, updatePhys=False, dist00=True, evalStress=False, reorderEvery=1000,
                     precision="fp64", geoFunctors=lambda: [], phyFunctors=lambda: [], lawFunctors=lambda: [])

    def __init__(self, geo=None, phy=None, law=None, **kw):
        super().__init__(**kw)
        if geo is not None:
            self.geoFunctors, self.phyFunctors, self.lawFunctors = list(geo), list(phy), list(law)


class PeriodicEngine(Engine):
    """PeriodicEngine attributes and its activation rule (core/Engine.hpp:129-153,
    core/Engine.cpp:169-195); realPeriod (wall-clock time) is not supported on this path."""
    _periodic = dict(stepPeriod=1, virtPeriod=0., realPeriod=0., stepModulo=True, nDo=-1, nDone=0, initRun=True,
                     virtLast=float("nan"), virtPrev=-1., stepLast=-1, stepPrev=-1)

    def _activate(self, stepNow, virtNow):
        """PeriodicEngine::isActivated at step `stepNow` (with its side effects)"""
        if self.realPeriod > 0:
            raise NotImplementedError("%s.realPeriod is not supported on this path" % type(self).__name__)
        if self.virtPeriod > 0 and virtNow is None:
            raise NotImplementedError("%s.virtPeriod is not supported here (use stepPeriod)" % type(self).__name__)
        initNow = self.stepLast < 0
        if initNow:
            self.virtLast, self.stepLast = virtNow, stepNow
        if (self.nDo < 0 or self.nDone < self.nDo) and (
                (self.virtPeriod > 0 and virtNow - self.virtLast >= self.virtPeriod) or
                (self.stepPeriod > 0 and ((stepNow - self.stepLast >= self.stepPeriod) or
                                          (self.stepModulo and stepNow % self.stepPeriod == 0)))):
            self.virtPrev, self.virtLast = self.virtLast, virtNow
            self.stepPrev, self.stepLast = self.stepLast, stepNow
            if initNow and not self.initRun:
                return False
            self.nDone += 1
            return True
        return False


class BoxOutlet(PeriodicEngine):
    """BoxOutlet (pkg/dem/Outlet.hpp:52-60) with the Outlet attributes (Outlet.hpp:22-38) and
    Outlet::run (Outlet.cpp:12-110) for uninodal particles: deletes (markMask 0) or marks the
    particles outside `box` (inside it with inside=True) matching `mask`.  It runs after the
    engines on the device; Scene.run pauses there at the steps it fires on.  The particles'
    kinetic energy goes to Scene.energy['kinOutlet'] when Scene.trackEnergy is set."""
    _defaults = dict(box=((0., 0., 0.), (0., 0., 0.)), inside=False, mask=0b0100, markMask=0, save=False, num=0, mass=0.,
                     currRate=float("nan"), currRateSmooth=1., diamMassTime=lambda: [], **PeriodicEngine._periodic)

    def _run(self, scene, step, time, dt):
        ctx, sid, a = scene._ctx, scene._sid, scene.dem._arrays
        lo, hi = np.asarray(self.box[0], np.float64), np.asarray(self.box[1], np.float64)
        ids = ctx.select_box(sid, lo, hi, self.inside, self.mask, self.markMask)
        deleting = self.markMask == 0
        m = a.mass[ids] if len(ids) else np.zeros(0)
        if deleting and scene.trackEnergy and len(ids):
            nd = ctx.download_nodes(sid, a.n_node)
            v, w = nd["vel"][ids], nd["angVel"][ids]
            ek = .5 * m * np.einsum("ij,ij->i", v, v) + .5 * np.einsum("ij,ij->i", w, a.inertia[ids] * w)
            scene._host_energy["kinOutlet"] = scene._host_energy.get("kinOutlet", 0.) + float(np.sum(ek))
        stepMass = float(np.sum(m))
        self.num += len(ids)
        self.mass += stepMass
        if self.save:
            for i, mi in zip(ids, m):
                self.diamMassTime.append((2 * float(a.radius[i]), float(mi), time))
        if len(ids):
            if deleting:
                ctx.remove_particles(sid, ids)
                scene.dem.par._removed(ids)
                scene.dem._mark_removed(ids)
            else:
                ctx.set_masks(sid, ids, a.mask[ids] | self.markMask)
                a.mask[ids] |= self.markMask
        span = self.stepPeriod if (self.stepPrev < 0 and self.stepPeriod > 0) else step - self.stepPrev
        den = span * dt
        # the reference divides in C++: 0/0 is NaN (the first run, fired at the step counting starts)
        rate = stepMass / den if den != 0 else (math.nan if stepMass == 0 else math.copysign(math.inf, stepMass))
        if math.isnan(self.currRate) or self.stepPrev < 0:
            self.currRate = rate
        else:
            self.currRate = (1 - self.currRateSmooth) * self.currRate + self.currRateSmooth * rate
        return ids


class DynDt(Engine):
    """pkg/dem/DynDt.hpp:20-26 with the PeriodicEngine attributes and state it runs by
    (core/Engine.hpp:141-153), placed after ContactLoop.  The state attributes (nDone,
    stepLast, stepPrev, virtLast, virtPrev) and `dt` (the dryRun result) are updated by
    Scene.run.  realPeriod (wall-clock time) is not supported on this path."""
    _defaults = dict(maxRelInc=1e-4, dryRun=False, dt=float("nan"), stepPeriod=1, virtPeriod=0., realPeriod=0.,
                     stepModulo=True, nDo=-1, nDone=0, initRun=True, virtLast=float("nan"), virtPrev=-1., stepLast=-1,
                     stepPrev=-1)
    _state = ("nDone", "stepLast", "stepPrev", "virtLast", "virtPrev", "dt")

    def _native(self, dtSafety):
        from ._native import make_dyndt
        if self.realPeriod > 0:
            raise NotImplementedError("DynDt.realPeriod (wall-clock periodicity) is not supported on this path")
        if 1. + self.maxRelInc == 1.:
            raise ValueError("DynDt: maxRelInc too small (1.0+maxRelInc==1.0)")  # DynDt::postLoad, DynDt.cpp:12-14
        return make_dyndt(maxRelInc=self.maxRelInc, dtSafety=dtSafety, virtPeriod=self.virtPeriod, virtLast=self.virtLast,
                          virtPrev=self.virtPrev, dt=self.dt, stepPeriod=self.stepPeriod, nDo=self.nDo, nDone=self.nDone,
                          stepLast=self.stepLast, stepPrev=self.stepPrev, enabled=0 if self.dead else 1,
                          stepModulo=int(bool(self.stepModulo)), initRun=int(bool(self.initRun)), dryRun=int(bool(self.dryRun)))

    def _update(self, d):
        for k in self._state:
            setattr(self, k, getattr(d, k))


# convenience: module-level constructor of a bare field with particles
def _mk_field(**kw):
    return DemField(**kw)
