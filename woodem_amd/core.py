"""Scene, Cell and EnergyTracker (module woo.core of the reference) driving the HIP path.

Scene.run(n) performs n Scene::doOneStep (core/Scene.cpp:319-424) through the C
ABI: prologue (cell, energy reset), Leapfrog, InsertionSortCollider,
ContactLoop, epilogue.  Engines must be the reference's defaultEngines order
(py/utils.py:76-83): Leapfrog(reset=True), InsertionSortCollider, ContactLoop.
"""
import math

import numpy as np

from . import dem as _dem
from ._native import (ENERGY_NAMES, LAW_HERTZ_DMT, LAW_IDEALELPL, PREC_FP64, PREC_MIXED, Context, WdmParams)

_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class Cell(_dem._Attrs):
    """Periodic cell, core/Cell.hpp:116-160 (homoDeform 4 = HOMO_GRADV2, the default)."""
    _defaults = dict(hSize=lambda: np.eye(3), gradV=lambda: np.zeros((3, 3)), nextGradV=lambda: np.zeros((3, 3)),
                     homoDeform=4)
    HomoNone, HomoPos, HomoVel, HomoVel2, HomoGradV2 = 0, 1, 2, 3, 4

    def setBox(self, *size):
        if len(size) == 1:
            size = size[0]
        self.hSize = np.diag(np.asarray(size, dtype=np.float64))

    @property
    def size(self):
        return np.linalg.norm(np.asarray(self.hSize), axis=0)


class EnergyTracker:
    """EnergyTracker, core/EnergyTracker.hpp:10-84 (read-only view after a run)."""

    def __init__(self, values=None, created=None, extra=None):
        self._v = dict()
        if values is not None:
            for k, name in enumerate(ENERGY_NAMES):
                if created[k]:
                    self._v[name] = float(values[k])
        for name, v in (extra or {}).items():  # terms added by engines on the host (Outlet's kinOutlet)
            self._v[name] = float(v)

    def __getitem__(self, k):
        return self._v[k]

    def __contains__(self, k):
        return k in self._v

    def keys(self):
        return list(self._v.keys())

    def items(self):
        return list(self._v.items())

    def __len__(self):
        return len(self._v)

    def total(self):
        return float(sum(self._v.values()))

    def relErr(self):
        s = sum(abs(v) for v in self._v.values())
        return float(sum(self._v.values()) / s) if s > 0 else 0.


class Scene(_dem._Attrs):
    """Scene, core/Scene.hpp:27-225."""
    _defaults = dict(dt=float("nan"), dtSafety=.9, trackEnergy=False, periodic=False, engines=lambda: [],
                     fields=lambda: [], cell=lambda: Cell())

    def __init__(self, **kw):
        super().__init__(**kw)
        self._ctx = None
        self._sid = None
        self._uploaded_sig = None
        self._stale = True
        self._con_cache = None
        self._host_energy = {}

    @property
    def dem(self):
        for f in self.fields:
            if isinstance(f, _dem.DemField):
                return f
        raise AttributeError("no DemField in Scene.fields")

    @property
    def step(self):
        if self._sid is None:
            return 0
        return self._ctx.stats(self._sid)["step"]

    # ---- engine attributes -> wdm_params ----
    def _params(self):
        eng = [e for e in self.engines if not getattr(e, "dead", False)]
        kinds = [type(e) for e in eng]
        # engines after ContactLoop: DynDt (on the device, inside each step), then host-side
        # engines (BoxOutlet) that Scene.run pauses for
        tail = []
        while eng and isinstance(eng[-1], (_dem.DynDt, _dem.BoxOutlet)):
            tail.insert(0, eng.pop())
            kinds.pop()
        dyn = [e for e in tail if isinstance(e, _dem.DynDt)]
        if len(dyn) > 1 or (dyn and tail[0] is not dyn[0]):
            raise ValueError("DynDt must follow ContactLoop directly (before any Outlet)")
        for e in tail:
            if isinstance(e, _dem.BoxOutlet) and (e.realPeriod > 0 or e.virtPeriod > 0):
                raise NotImplementedError("BoxOutlet runs by stepPeriod on this path")
        if len(eng) != 3 or not (isinstance(eng[0], _dem.Leapfrog) and isinstance(eng[1], _dem.InsertionSortCollider)
                                 and isinstance(eng[2], _dem.ContactLoop)):
            raise ValueError("engines must be [Leapfrog, InsertionSortCollider, ContactLoop] (got %s)" % [k.__name__ for k in kinds])
        lf, isc, cl = eng
        if not lf.reset:
            raise ValueError("Leapfrog.reset must be True (no ForceResetter on this path)")
        if not cl.applyForces or cl.evalStress or cl.updatePhys:
            raise ValueError("ContactLoop: applyForces=True, evalStress=False, updatePhys=False are required")
        bo1 = [f for f in isc.boundFunctors if isinstance(f, _dem.Bo1_Sphere_Aabb)]
        cgs = [f for f in cl.geoFunctors if isinstance(f, _dem.Cg2_Sphere_Sphere_L6Geom)]
        df = bo1[0].distFactor if bo1 else 1.
        if cgs and cgs[0].distFactor != df:
            raise ValueError("Bo1_Sphere_Aabb.distFactor must equal Cg2_Sphere_Sphere_L6Geom.distFactor")
        for f in cl.geoFunctors:
            if isinstance(f, _dem.Cg2_Any_Any_L6Geom__Base) and (f.noRatch or f.approxMask):
                raise ValueError("noRatch/approxMask must keep their defaults on this path")
        renorm = [f.trsfRenorm for f in cl.geoFunctors if isinstance(f, _dem.Cg2_Any_Any_L6Geom__Base)]
        cp2 = cl.phyFunctors[0]
        law = cl.lawFunctors[0]
        p = WdmParams()
        S_dem = self.dem
        p.dt = self.dt
        p.damping = lf.damping
        for k in range(3):
            p.gravity[k] = float(S_dem.gravity[k])
        p.verletDist = isc.verletDist
        p.distFactor = df
        p.trackEnergy = int(bool(self.trackEnergy))
        p.trsfRenorm = int(renorm[0]) if renorm else 100
        p.dist00 = int(bool(cl.dist00))
        p.loneMask = int(S_dem.loneMask)
        p.precision = PREC_MIXED if cl.precision == "mixed" else PREC_FP64
        if isinstance(cp2, _dem.Cp2_FrictMat_HertzPhys):
            if not isinstance(law, _dem.Law2_L6Geom_HertzPhys_DMT):
                raise ValueError("Cp2_FrictMat_HertzPhys needs Law2_L6Geom_HertzPhys_DMT")
            p.law = LAW_HERTZ_DMT
            p.poisson, p.en, p.gamma, p.alpha = cp2.poisson, cp2.en, cp2.gamma, cp2.alpha
            p.noAttraction = int(bool(law.noAttraction))
        elif isinstance(cp2, _dem.Cp2_FrictMat_FrictPhys):
            if not isinstance(law, _dem.Law2_L6Geom_FrictPhys_IdealElPl):
                raise ValueError("Cp2_FrictMat_FrictPhys needs Law2_L6Geom_FrictPhys_IdealElPl")
            if law.iniEqlb or law.relRollStiff > 0:
                raise ValueError("iniEqlb and rolling resistance are outside this path")
            p.law = LAW_IDEALELPL
            p.noSlip, p.noBreak, p.noFrict = int(law.noSlip), int(law.noBreak), int(law.noFrict)
            p.poisson, p.en = .2, float("nan")
        else:
            raise ValueError("unsupported Cp2 functor %s" % type(cp2).__name__)
        if cp2.tanPhi is not None:
            raise ValueError("MatchMaker tanPhi is outside this path")
        if math.isnan(p.dt) or math.isinf(p.dt) or p.dt < 0:
            # Scene::doOneStep auto-dt: dtSafety * min critDt of the engines and fields
            # (core/Scene.cpp:333-355): DemField's p-wave step, DynDt's stiffness step once
            # there are contacts
            crit = S_dem.critDt()
            if dyn and self._sid is not None:
                crit = min(crit, self._ctx.dyndt_crit(self._sid))
            p.dt = self.dtSafety * crit
            self.dt = p.dt
        return p

    def _host_engines(self):
        return [e for e in self.engines if isinstance(e, _dem.BoxOutlet) and not getattr(e, "dead", False)]

    def _dyndt_engine(self):
        d = [e for e in self.engines if isinstance(e, _dem.DynDt) and not getattr(e, "dead", False)]
        return d[0] if d else None

    def _set_dyndt(self):
        from ._native import make_dyndt
        e = self._dyndt_engine()
        self._ctx.set_dyndt(self._sid, e._native(self.dtSafety) if e else make_dyndt(enabled=0))

    def _after_run(self):
        """Scene.dt/time and the DynDt engine's state after steps on the device."""
        t, dt = self._ctx.time(self._sid)
        self._time, self.dt = t, dt
        e = self._dyndt_engine()
        if e is not None:
            e._update(self._ctx.dyndt(self._sid))

    @property
    def time(self):
        return getattr(self, "_time", 0.)

    def _ensure(self):
        f = self.dem
        if self._ctx is None:
            self._ctx = default_context()
        if f._dirty or self._sid is None:
            arr = f._compile()
            if self._sid is not None:
                self._ctx.scene_free(self._sid)
            self._sid = self._ctx.scene_alloc(arr.n_par, arr.n_node, arr.n_mat)
            self._ctx.upload(self._sid, arr)
            f._dirty = False
            f._scene = self
        c = self.cell
        self._ctx.set_cell(self._sid, self.periodic, c.hSize, c.gradV, c.nextGradV, c.homoDeform)
        self._ctx.set_params(self._sid, self._params())
        self._set_dyndt()

    def run(self, nSteps=1, wait=True):
        first = self._sid is None or self.dem._dirty
        if first:
            self._ensure()
        else:
            self._ctx.set_params(self._sid, self._params())
            self._set_dyndt()
        host = self._host_engines()
        if not host:
            self._ctx.step([self._sid], int(nSteps))
        else:
            # step on the device up to each step a host engine fires on, then run it there (the
            # reference runs it after ContactLoop of that step; the epilogue in between changes
            # no state it reads)
            left = int(nSteps)
            while left > 0:
                s0 = self._ctx.stats(self._sid)["step"]
                fire, fired = None, []
                for j in range(left):
                    fired = [e for e in host if e._activate(s0 + j, None)]
                    if fired:
                        fire = j
                        break
                if fire is None:
                    self._ctx.step([self._sid], left)
                    break
                if fire > 0:
                    self._ctx.step([self._sid], fire)
                t, dt = self._ctx.time(self._sid)
                self._ctx.step([self._sid], 1)
                for e in fired:
                    e._run(self, s0 + fire, t, dt)
                left -= fire + 1
        self._after_run()
        self._stale = True
        self._con_cache = None
        h, g = self._ctx.cell(self._sid)
        self.cell.hSize, self.cell.gradV = h, g

    def one(self):
        self.run(1)

    def wait(self):
        if self._ctx is not None:
            self._ctx.sync()

    def _download_if_stale(self):
        if self._sid is None:
            return
        if self._stale:
            f = self.dem
            f._arrays_cur = self._ctx.download_nodes(self._sid, f._arrays.n_node)
            self._stale = False

    def _contacts(self):
        if self._con_cache is None:
            self._con_cache = self._ctx.contacts(self._sid)
        return self._con_cache

    @property
    def energy(self):
        if self._sid is None:
            return EnergyTracker()
        v, c = self._ctx.energy(self._sid)
        return EnergyTracker(v, c, self._host_energy)

    def stats(self):
        return self._ctx.stats(self._sid)
