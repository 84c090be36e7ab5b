"""ctypes binding of libwoodem_hip.so (the C ABI declared in include/woodem_hip.h).

The library is the only compute path: if it is missing, importing this module
raises, and there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwoodem_hip.so")

WDM_OK = 0
ERRORS = {-1: "WDM_E_ARG", -2: "WDM_E_HIP", -3: "WDM_E_OOM", -4: "WDM_E_NAN", -5: "WDM_E_OVERSIZE",
          -6: "WDM_E_DEGENERATE", -7: "WDM_E_UNSUPPORTED", -8: "WDM_E_STATE"}

SHAPE_NONE, SHAPE_SPHERE, SHAPE_WALL, SHAPE_FACET = 0, 1, 2, 3
LAW_IDEALELPL, LAW_HERTZ_DMT = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
ENERGY_NAMES = ["kinetic", "elast", "plast", "grav", "nonviscDamp", "viscN", "viscT", "broken"]


class WdmParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("damping", C.c_double), ("gravity", C.c_double * 3), ("verletDist", C.c_double),
        ("distFactor", C.c_double), ("poisson", C.c_double), ("en", C.c_double), ("gamma", C.c_double),
        ("alpha", C.c_double),
        ("law", C.c_int32), ("noSlip", C.c_int32), ("noBreak", C.c_int32), ("noFrict", C.c_int32),
        ("noAttraction", C.c_int32), ("trackEnergy", C.c_int32), ("trsfRenorm", C.c_int32), ("dist00", C.c_int32),
        ("loneMask", C.c_uint32), ("precision", C.c_int32),
    ]


class WdmDynDt(C.Structure):
    """wdm_dyndt: DynDt attributes with its PeriodicEngine attributes and state."""
    _fields_ = [
        ("maxRelInc", C.c_double), ("dtSafety", C.c_double), ("virtPeriod", C.c_double), ("virtLast", C.c_double),
        ("virtPrev", C.c_double), ("dt", C.c_double),
        ("stepPeriod", C.c_int64), ("nDo", C.c_int64), ("nDone", C.c_int64), ("stepLast", C.c_int64), ("stepPrev", C.c_int64),
        ("enabled", C.c_int32), ("stepModulo", C.c_int32), ("initRun", C.c_int32), ("dryRun", C.c_int32),
    ]


DYNDT_FIELDS = [f for f, _ in WdmDynDt._fields_]


def make_dyndt(**kw):
    """WdmDynDt with the reference's defaults (DynDt.hpp:22-24, core/Engine.hpp:141-153)
    overridden by keyword arguments."""
    d = WdmDynDt()
    defaults = dict(maxRelInc=1e-4, dtSafety=.9, virtPeriod=0., virtLast=float("nan"), virtPrev=-1., dt=float("nan"),
                    stepPeriod=1, nDo=-1, nDone=0, stepLast=-1, stepPrev=-1, enabled=1, stepModulo=1, initRun=1, dryRun=0)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(d, k, v)
    return d


class WdmError(RuntimeError):
    """Error returned by the C ABI; `code` is the WDM_E_* value."""

    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libwoodem_hip.so is not built (run __graft_entry__.build() or python woodem_amd/build.py)")
    L = C.CDLL(LIB_PATH)
    P, I32, I64, D, U8, U32 = C.c_void_p, C.c_int32, C.c_int64, C.c_double, C.c_uint8, C.c_uint32
    pd, pi, pu, pl = C.POINTER(D), C.POINTER(I32), C.POINTER(U32), C.POINTER(I64)
    sig = {
        "wdm_create": (C.c_int, [C.POINTER(P), C.c_int, C.c_size_t]),
        "wdm_destroy": (None, [P]),
        "wdm_last_error": (C.c_char_p, [P]),
        "wdm_scene_alloc": (C.c_int, [P, I32, I32, I32, pi]),
        "wdm_scene_free": (C.c_int, [P, I32]),
        "wdm_scene_footprint": (C.c_size_t, [I32, I32, D]),
        "wdm_upload_nodes": (C.c_int, [P, I32, I32, pd, pd, pd, pd, pd, pd, pu, pd, pd]),
        "wdm_upload_particles": (C.c_int, [P, I32, I32, pi, pi, pd, pi, pi, pu, pi]),
        "wdm_upload_materials": (C.c_int, [P, I32, I32, pd, pd, pd, pd]),
        "wdm_upload_clumps": (C.c_int, [P, I32, I32, pi, pi, pi, pd, pd]),
        "wdm_set_cell": (C.c_int, [P, I32, I32, pd, pd, pd, I32]),
        "wdm_set_params": (C.c_int, [P, I32, C.POINTER(WdmParams)]),
        "wdm_step_begin": (C.c_int, [P, I32]),
        "wdm_leapfrog_run": (C.c_int, [P, I32]),
        "wdm_collider_run": (C.c_int, [P, I32]),
        "wdm_contact_loop_run": (C.c_int, [P, I32]),
        "wdm_step_end": (C.c_int, [P, I32]),
        "wdm_scene_error": (C.c_int, [P, I32, pi]),
        "wdm_step": (C.c_int, [P, pi, I32, I32]),
        "wdm_download_nodes": (C.c_int, [P, I32, pd, pd, pd, pd, pd, pd]),
        "wdm_num_contacts": (C.c_int, [P, I32, pl]),
        "wdm_contact_work": (C.c_int, [P, I32, pl, pl]),
        "wdm_get_contacts": (C.c_int, [P, I32, I64, pi, pi, pi, C.POINTER(U8), pd, pd, pd, pd, pl]),
        "wdm_get_energy": (C.c_int, [P, I32, pd, pi]),
        "wdm_get_stats": (C.c_int, [P, I32, pl, pl, pd]),
        "wdm_get_cell": (C.c_int, [P, I32, pd, pd]),
        "wdm_sync": (C.c_int, [P]),
        "wdm_kernel_times": (C.c_int, [P, I32, C.c_char_p, pd, pl, pi]),
        "wdm_enable_timing": (C.c_int, [P, I32]),
        "wdm_set_tuning": (C.c_int, [P, I32, I32, D]),
        "wdm_set_broadphase": (C.c_int, [P, I32]),
        "wdm_get_broadphase": (C.c_int, [P, pi]),
        "wdm_set_relayout": (C.c_int, [P, I32, I64]),
        "wdm_get_relayout": (C.c_int, [P, I32, pi, pl, pl]),
        "wdm_get_tuning": (C.c_int, [P, pi, pi, pd, pl, pl, pl]),
        "wdm_relax_overlaps": (C.c_int, [P, I32, I32]),
        "wdm_set_dyndt": (C.c_int, [P, I32, C.POINTER(WdmDynDt)]),
        "wdm_get_dyndt": (C.c_int, [P, I32, C.POINTER(WdmDynDt)]),
        "wdm_dyndt_crit": (C.c_int, [P, I32, pd]),
        "wdm_dyndt_run": (C.c_int, [P, I32]),
        "wdm_get_time": (C.c_int, [P, I32, pd, pd]),
        "wdm_get_contact_state": (C.c_int, [P, I32, I64, pd, pd, pl]),
        "wdm_upload_contacts": (C.c_int, [P, I32, I64, pi, pi, pi, pd, pd, pd, pd, pd]),
        "wdm_set_energy": (C.c_int, [P, I32, pd, pi]),
        "wdm_set_time": (C.c_int, [P, I32, I64, D]),
        "wdm_remove_particles": (C.c_int, [P, I32, I32, pi]),
        "wdm_insert_spheres": (C.c_int, [P, I32, I32, pd, pd, pd, pd, pd, pd, pi, pu, pi]),
        "wdm_get_free_ids": (C.c_int, [P, I32, I32, pi, pi, pi]),
        "wdm_set_free_ids": (C.c_int, [P, I32, I32, pi, I32]),
        "wdm_set_masks": (C.c_int, [P, I32, I32, pi, pu]),
        "wdm_select_box": (C.c_int, [P, I32, pd, pd, I32, U32, U32, I32, pi, pi]),
        "wdm_select_box_noalloc": (C.c_int, [P, I32, pd, pd, I32, U32, U32, I32, pi, pi]),
        "wdm_make_cloud": (C.c_int, [P, pd, pd, D, D, I32, I32, D, pd, pd, I32, I32, pu, I64, pd, pd, pl, pd]),
        "wdm_regular_pack": (C.c_int, [P, I32, I32, pd, D, D, I64, pd, pl]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def ptr(a, ctype):
    """Pointer to a contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


class Context:
    """One wdm_ctx: a HIP stream on one device plus its scenes."""

    def __init__(self, device=0, hbm_budget=0):
        L = lib()
        h = C.c_void_p()
        rc = L.wdm_create(C.byref(h), device, hbm_budget)
        if rc != WDM_OK:
            raise WdmError(rc, "wdm_create failed (no HIP device?)")
        self.h = h
        self.L = L

    def check(self, rc):
        if rc != WDM_OK:
            raise WdmError(rc, self.L.wdm_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.L.wdm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- thin wrappers ----
    def scene_alloc(self, n_par, n_node, n_mat):
        sid = C.c_int32()
        self.check(self.L.wdm_scene_alloc(self.h, n_par, n_node, n_mat, C.byref(sid)))
        return sid.value

    def scene_free(self, sid):
        self.check(self.L.wdm_scene_free(self.h, sid))

    def upload(self, sid, arrs):
        """Upload a DemArrays (woodem_amd.soa) into scene `sid`."""
        a = arrs
        L, h = self.L, self.h
        D, I32, U32 = C.c_double, C.c_int32, C.c_uint32
        self.check(L.wdm_upload_nodes(h, sid, a.n_node, ptr(a.pos, D), ptr(a.ori, D), ptr(a.vel, D), ptr(a.angVel, D),
                                      ptr(a.mass, D), ptr(a.inertia, D), ptr(a.flags, U32), ptr(a.force, D), ptr(a.torque, D)))
        self.check(L.wdm_upload_particles(h, sid, a.n_par, ptr(a.shape, I32), ptr(a.nodes, I32), ptr(a.radius, D),
                                          ptr(a.axis, I32), ptr(a.sense, I32), ptr(a.mask, U32), ptr(a.mat, I32)))
        self.check(L.wdm_upload_materials(h, sid, a.n_mat, ptr(a.density, D), ptr(a.young, D), ptr(a.ktDivKn, D),
                                          ptr(a.tanPhi, D)))
        c = getattr(a, "clumps", None)
        if c:
            self.check(L.wdm_upload_clumps(h, sid, len(c["master"]), ptr(c["master"], I32), ptr(c["off"], I32),
                                           ptr(c["members"], I32), ptr(c["relPos"], D), ptr(c["relOri"], D)))

    def set_cell(self, sid, periodic, hSize, gradV, nextGradV, homoDeform):
        D = C.c_double
        h9 = np.ascontiguousarray(hSize, dtype=np.float64).reshape(9)
        g9 = np.ascontiguousarray(gradV, dtype=np.float64).reshape(9)
        n9 = np.ascontiguousarray(nextGradV, dtype=np.float64).reshape(9)
        self.check(self.L.wdm_set_cell(self.h, sid, int(periodic), ptr(h9, D), ptr(g9, D), ptr(n9, D), int(homoDeform)))

    def set_params(self, sid, p):
        self.check(self.L.wdm_set_params(self.h, sid, C.byref(p)))

    def step(self, sids, n):
        """n whole steps of every listed scene (Scene::doOneStep).  A scene stopped by a device
        error does not keep the others from their steps; the call then raises (scene_error
        names the stopped scenes)."""
        arr = np.ascontiguousarray(sids, dtype=np.int32)
        self.check(self.L.wdm_step(self.h, ptr(arr, C.c_int32), len(arr), int(n)))

    def scene_error(self, sid):
        """(WDM_E_* code, message) of the device error that stopped scene `sid`; (0, "") while it runs"""
        code = C.c_int32()
        self.check(self.L.wdm_scene_error(self.h, sid, C.byref(code)))
        return code.value, (self.L.wdm_last_error(self.h).decode() if code.value else "")

    # ---- one engine at a time: the calls a host Scene::doOneStep makes (core/Scene.cpp:319-424) ----
    def step_begin(self, sid):
        """prologue: cell update, EnergyTracker resettables"""
        self.check(self.L.wdm_step_begin(self.h, sid))

    def leapfrog_run(self, sid):
        """Leapfrog::run (pkg/dem/Leapfrog.cpp:103-255)"""
        self.check(self.L.wdm_leapfrog_run(self.h, sid))

    def collider_run(self, sid):
        """InsertionSortCollider::run (pkg/dem/InsertionSortCollider.cpp:445-646)"""
        self.check(self.L.wdm_collider_run(self.h, sid))

    def contact_loop_run(self, sid):
        """ContactLoop::run (pkg/dem/ContactLoop.cpp:93-259)"""
        self.check(self.L.wdm_contact_loop_run(self.h, sid))

    def dyndt_engine_run(self, sid):
        """DynDt after ContactLoop, when its PeriodicEngine fires at this step"""
        self.check(self.L.wdm_dyndt_run(self.h, sid))

    def step_end(self, sid):
        """epilogue: step += 1, time += dt, next dt"""
        self.check(self.L.wdm_step_end(self.h, sid))

    def download_nodes(self, sid, n_node):
        D = C.c_double
        out = {k: np.empty((n_node, 3)) for k in ("pos", "vel", "angVel", "force", "torque")}
        out["ori"] = np.empty((n_node, 4))
        self.check(self.L.wdm_download_nodes(self.h, sid, ptr(out["pos"], D), ptr(out["ori"], D), ptr(out["vel"], D),
                                             ptr(out["angVel"], D), ptr(out["force"], D), ptr(out["torque"], D)))
        return out

    def contacts(self, sid):
        n = C.c_int64()
        self.check(self.L.wdm_num_contacts(self.h, sid, C.byref(n)))
        cap = n.value
        D, I32 = C.c_double, C.c_int32
        idA = np.empty(cap, np.int32)
        idB = np.empty(cap, np.int32)
        cd = np.empty((cap, 3), np.int32)
        real = np.empty(cap, np.uint8)
        fl = np.empty((cap, 3))
        cp = np.empty((cap, 3))
        tr = np.empty((cap, 9))
        un = np.empty(cap)
        nout = C.c_int64()
        self.check(self.L.wdm_get_contacts(self.h, sid, cap, ptr(idA, I32), ptr(idB, I32), ptr(cd, I32),
                                           real.ctypes.data_as(C.POINTER(C.c_uint8)), ptr(fl, D), ptr(cp, D), ptr(tr, D),
                                           ptr(un, D), C.byref(nout)))
        k = nout.value
        return dict(idA=idA[:k], idB=idB[:k], cellDist=cd[:k], isReal=real[:k].astype(bool), forceLocal=fl[:k],
                    contPt=cp[:k], trsf=tr[:k].reshape(-1, 3, 3), uN=un[:k])

    def contact_work(self, sid):
        """(contacts the ContactLoop processes, state rows) at the last full run"""
        a, r = C.c_int64(), C.c_int64()
        self.check(self.L.wdm_contact_work(self.h, sid, C.byref(a), C.byref(r)))
        return a.value, r.value

    def energy(self, sid):
        v = np.zeros(8)
        cr = np.zeros(8, np.int32)
        self.check(self.L.wdm_get_energy(self.h, sid, ptr(v, C.c_double), ptr(cr, C.c_int32)))
        return v, cr.astype(bool)

    def stats(self, sid):
        s, f, vd = C.c_int64(), C.c_int64(), C.c_double()
        self.check(self.L.wdm_get_stats(self.h, sid, C.byref(s), C.byref(f), C.byref(vd)))
        return dict(step=s.value, nFullRuns=f.value, verletDist=vd.value)

    def cell(self, sid):
        h = np.zeros(9)
        g = np.zeros(9)
        self.check(self.L.wdm_get_cell(self.h, sid, ptr(h, C.c_double), ptr(g, C.c_double)))
        return h.reshape(3, 3), g.reshape(3, 3)

    def sync(self):
        self.check(self.L.wdm_sync(self.h))

    def set_tuning(self, window=None, graphs=None, cell_factor=None):
        """Launch-shape knobs of step(): steps per window (1 = step by step), captured
        graphs on/off, collider cell width factor (>= 1); None keeps the current value.
        Results do not depend on them."""
        t = self.tuning()
        w = t["window"] if window is None else int(window)
        g = t["graphs"] if graphs is None else int(bool(graphs))
        f = t["cell_factor"] if cell_factor is None else float(cell_factor)
        self.check(self.L.wdm_set_tuning(self.h, w, g, f))

    def set_broadphase(self, mode):
        """Collider broadphase: 0 the multi-level grid (default), 1 sort and sweep; the contact
        set is the same."""
        self.check(self.L.wdm_set_broadphase(self.h, int(mode)))

    def broadphase(self):
        m = C.c_int32()
        self.check(self.L.wdm_get_broadphase(self.h, C.byref(m)))
        return m.value

    def set_relayout(self, steps, min_contacts=1 << 20):
        """Steps between relayouts of the ContactLoop work list of a scene with at least
        min_contacts contacts, after its last full run or relayout (0: never); memory order
        only, results do not change."""
        self.check(self.L.wdm_set_relayout(self.h, int(steps), int(min_contacts)))

    def relayout(self, sid=-1):
        """(steps between relayouts, min_contacts, relayouts of scene sid so far)"""
        st, mn, n = C.c_int32(), C.c_int64(), C.c_int64()
        self.check(self.L.wdm_get_relayout(self.h, int(sid), C.byref(st), C.byref(mn), C.byref(n)))
        return st.value, mn.value, n.value

    def tuning(self):
        w, g = C.c_int32(), C.c_int32()
        f = C.c_double()
        nw, nc, nl = C.c_int64(), C.c_int64(), C.c_int64()
        self.check(self.L.wdm_get_tuning(self.h, C.byref(w), C.byref(g), C.byref(f), C.byref(nw), C.byref(nc), C.byref(nl)))
        return dict(window=w.value, graphs=g.value, cell_factor=f.value, windows=nw.value, graph_captures=nc.value,
                    graph_launches=nl.value)

    def enable_timing(self, on=True):
        self.check(self.L.wdm_enable_timing(self.h, int(on)))

    def kernel_times(self):
        cap = 64
        names = C.create_string_buffer(64 * cap)
        ms = np.zeros(cap)
        la = np.zeros(cap, np.int64)
        n = C.c_int32()
        self.check(self.L.wdm_kernel_times(self.h, cap, names, ptr(ms, C.c_double), ptr(la, C.c_int64), C.byref(n)))
        out = {}
        raw = names.raw
        for i in range(n.value):
            nm = raw[64 * i:64 * i + 64].split(b"\0", 1)[0].decode()
            out[nm] = (float(ms[i]), int(la[i]))
        return out

    def set_dyndt(self, sid, d):
        """DynDt engine of scene `sid` (a WdmDynDt, see make_dyndt; enabled=0 removes it)."""
        self.check(self.L.wdm_set_dyndt(self.h, sid, C.byref(d)))

    def dyndt(self, sid):
        d = WdmDynDt()
        self.check(self.L.wdm_get_dyndt(self.h, sid, C.byref(d)))
        return d

    def dyndt_crit(self, sid):
        """DynDt::critDt_compute on the contacts of the last ContactLoop."""
        v = C.c_double()
        self.check(self.L.wdm_dyndt_crit(self.h, sid, C.byref(v)))
        return v.value

    def dyndt_run(self, sid):
        self.check(self.L.wdm_dyndt_run(self.h, sid))

    def time(self, sid):
        """(Scene.time, Scene.dt)"""
        t, d = C.c_double(), C.c_double()
        self.check(self.L.wdm_get_time(self.h, sid, C.byref(t), C.byref(d)))
        return t.value, d.value

    def contact_state(self, sid):
        """phys (5 values) and L6Geom.lens (2) of the contacts, in contacts() order"""
        n = C.c_int64()
        self.check(self.L.wdm_num_contacts(self.h, sid, C.byref(n)))
        ph = np.empty((n.value, 5))
        ln = np.empty((n.value, 2))
        k = C.c_int64()
        self.check(self.L.wdm_get_contact_state(self.h, sid, n.value, ptr(ph, C.c_double), ptr(ln, C.c_double), C.byref(k)))
        return ph[:k.value], ln[:k.value]

    def upload_contacts(self, sid, idA, idB, cellDist, trsf, forceLocal, phys, lens, contPt):
        """real contacts with their state (checkpoint restore)"""
        i32 = lambda x: np.ascontiguousarray(x, np.int32)
        f64 = lambda x: np.ascontiguousarray(x, np.float64)
        idA, idB, cd = i32(idA), i32(idB), i32(cellDist).reshape(-1, 3)
        tr, fl, ph, ln, cp = f64(trsf).reshape(-1, 9), f64(forceLocal), f64(phys), f64(lens), f64(contPt)
        n = len(idA)
        self.check(self.L.wdm_upload_contacts(self.h, sid, n, ptr(idA, C.c_int32), ptr(idB, C.c_int32), ptr(cd, C.c_int32),
                                              ptr(tr, C.c_double), ptr(fl, C.c_double), ptr(ph, C.c_double),
                                              ptr(ln, C.c_double), ptr(cp, C.c_double)))

    def set_energy(self, sid, values, created):
        v = np.ascontiguousarray(values, np.float64)
        c = np.ascontiguousarray(np.asarray(created).astype(np.int32))
        self.check(self.L.wdm_set_energy(self.h, sid, ptr(v, C.c_double), ptr(c, C.c_int32)))

    def set_time(self, sid, step, time):
        self.check(self.L.wdm_set_time(self.h, sid, int(step), float(time)))

    def remove_particles(self, sid, ids):
        """DemField::removeParticle of each id (the slot stays as a hole)"""
        ids = np.ascontiguousarray(ids, np.int32)
        self.check(self.L.wdm_remove_particles(self.h, sid, len(ids), ptr(ids, C.c_int32)))

    def insert_spheres(self, sid, pos, radius, mass, inertia, mat, mask, vel=None, angVel=None):
        """new spheres in free slots (ParticleContainer ids); returns their ids"""
        f64 = lambda x: None if x is None else np.ascontiguousarray(x, np.float64)
        pos, vel, angVel = f64(pos).reshape(-1, 3), f64(vel), f64(angVel)
        n = len(pos)
        bc = lambda x, t: np.ascontiguousarray(np.broadcast_to(np.asarray(x, t), (n,)))
        radius, mass, inertia = bc(radius, np.float64), bc(mass, np.float64), bc(inertia, np.float64)
        mat, mask = bc(mat, np.int32), bc(mask, np.uint32)
        ids = np.empty(n, np.int32)
        self.check(self.L.wdm_insert_spheres(self.h, sid, n, ptr(pos, C.c_double), ptr(vel, C.c_double), ptr(angVel, C.c_double),
                                             ptr(radius, C.c_double), ptr(mass, C.c_double), ptr(inertia, C.c_double),
                                             ptr(mat, C.c_int32), ptr(mask, C.c_uint32), ptr(ids, C.c_int32)))
        return ids

    def free_ids(self, sid):
        """(holes in reuse order, end of the particles in use)"""
        n, e = C.c_int32(), C.c_int32()
        self.check(self.L.wdm_get_free_ids(self.h, sid, 0, None, C.byref(n), C.byref(e)))
        ids = np.empty(n.value, np.int32)
        self.check(self.L.wdm_get_free_ids(self.h, sid, n.value, ptr(ids, C.c_int32), C.byref(n), C.byref(e)))
        return ids, e.value

    def set_free_ids(self, sid, ids, par_end):
        ids = np.ascontiguousarray(ids, np.int32)
        self.check(self.L.wdm_set_free_ids(self.h, sid, len(ids), ptr(ids, C.c_int32), int(par_end)))

    def set_masks(self, sid, ids, masks):
        ids = np.ascontiguousarray(ids, np.int32)
        masks = np.ascontiguousarray(np.broadcast_to(np.asarray(masks, np.uint32), ids.shape))
        self.check(self.L.wdm_set_masks(self.h, sid, len(ids), ptr(ids, C.c_int32), ptr(masks, C.c_uint32)))

    def select_box(self, sid, lo, hi, inside=False, mask=0, markMask=0):
        """ids Outlet::run would delete (or mark) for a BoxOutlet, ascending"""
        lo, hi = np.ascontiguousarray(lo, np.float64), np.ascontiguousarray(hi, np.float64)
        n = C.c_int32()
        self.check(self.L.wdm_select_box(self.h, sid, ptr(lo, C.c_double), ptr(hi, C.c_double), int(bool(inside)), int(mask),
                                         int(markMask), 0, None, C.byref(n)))
        ids = np.empty(n.value, np.int32)
        self.check(self.L.wdm_select_box(self.h, sid, ptr(lo, C.c_double), ptr(hi, C.c_double), int(bool(inside)), int(mask),
                                         int(markMask), n.value, ptr(ids, C.c_int32), C.byref(n)))
        return ids

    def select_box_noalloc(self, sid, lo, hi, inside=False, mask=0, markMask=0, cap=None):
        """select_box through wdm_select_box_noalloc (no device allocation; scenes of at most
        4,194,304 particles); returns (the first `cap` ids, how many were selected)"""
        lo, hi = np.ascontiguousarray(lo, np.float64), np.ascontiguousarray(hi, np.float64)
        cap = int(cap) if cap is not None else 1 << 22
        ids = np.empty(max(cap, 1), np.int32)
        n = C.c_int32()
        self.check(self.L.wdm_select_box_noalloc(self.h, sid, ptr(lo, C.c_double), ptr(hi, C.c_double), int(bool(inside)), int(mask),
                                                 int(markMask), cap, ptr(ids, C.c_int32), C.byref(n)))
        return ids[:min(n.value, cap)], n.value

    def make_cloud(self, mn, mx, rMean, rRelFuzz, num, periodic, porosity, psdSizes, psdCumm, distributeMass, state, cap):
        """SpherePack::makeCloud on the device; `state` (uint32 array of 1) is advanced in place.
        Returns (centres, radii, appliedPsdScaling)."""
        f64 = lambda x: np.ascontiguousarray(x, np.float64)
        mn, mx, ps, pc = f64(mn), f64(mx), f64(psdSizes), f64(psdCumm)
        c = np.empty((cap, 3))
        r = np.empty(cap)
        n = C.c_int64()
        sc = C.c_double()
        self.check(self.L.wdm_make_cloud(self.h, ptr(mn, C.c_double), ptr(mx, C.c_double), float(rMean), float(rRelFuzz), int(num),
                                         int(bool(periodic)), float(porosity), ptr(ps, C.c_double) if len(ps) else None,
                                         ptr(pc, C.c_double) if len(pc) else None, len(ps), int(bool(distributeMass)),
                                         ptr(state, C.c_uint32), int(cap), ptr(c, C.c_double), ptr(r, C.c_double), C.byref(n), C.byref(sc)))
        return c[:n.value], r[:n.value], sc.value

    def regular_pack(self, hexa, kind, params, radius, gap, cap=None):
        """pack.regularOrtho (hexa false) / regularHexa on the device for predicate `kind`
        (0 inAlignedBox, 1 inSphere, 2 inCylinder) with its parameters; returns the centres
        (all of them, or an error when there are more than `cap`)."""
        prm = np.ascontiguousarray(params, np.float64)
        n = C.c_int64()
        args = (int(bool(hexa)), int(kind), ptr(prm, C.c_double), float(radius), float(gap))
        self.check(self.L.wdm_regular_pack(self.h, *args, 0, None, C.byref(n)))
        m = n.value if cap is None else min(n.value, int(cap))
        c = np.empty((max(m, 1), 3))
        self.check(self.L.wdm_regular_pack(self.h, *args, m, ptr(c, C.c_double), C.byref(n)))
        return c[:n.value]

    def relax(self, sid, iters):
        self.check(self.L.wdm_relax_overlaps(self.h, sid, int(iters)))

    @staticmethod
    def footprint(n_par, n_node, contacts_per_particle):
        return lib().wdm_scene_footprint(n_par, n_node, contacts_per_particle)
