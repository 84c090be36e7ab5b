"""ctypes wrapper of the CPU oracle (oracle/liboracle_dem.so) -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.

Parity-pinning status of the oracle's functions (details in DESIGN.md):
  pinned by reference known answers / closed forms: pWaveDt, Sphere mass/inertia,
    bounds, contact set (vs brute force), sphere/wall/facet L6Geom creation,
    FrictPhys, IdealElPl slip cap, plastic dissipation (friction work per step) and
    elastic energy, Hertz collision time, restitution, Hertz/DMT elast (Hertz
    potential, also against the elastic force's work), viscN, viscT and plast (dashpot
    work, released shear-spring energy) step by step, collision energy balance,
    ballistic leapfrog, damping, rotation composition about non-parallel axes, GRADV2
    cell, incident velocity under PBC, energy balance, DynDt (single-contact stiffness
    closed forms, maxRelInc cap, dryRun, PeriodicEngine activation), checkpoint restore
    (orc_set_contacts: continues like the uninterrupted run; planted mistakes diverge);
  multi-step contact history by closed forms: L6Geom trsf of a spinning wall contact
    over 20,000 steps through the renormalisation, and the IdealElPl shear history of a
    sliding sphere's transition to rolling (tests/test_oracle.py);
  parity unpinned (no reference vector exists and no closed form fixes it): the summation
    order of forces and energies (Woo's OpenMP order is not deterministic; the oracle and
    the kernels sum in (min, max) key order, Woo's deterministic mode).
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle_dem.so")

ENERGY_NAMES = ["kinetic", "elast", "plast", "grav", "nonviscDamp", "viscN", "viscT", "broken"]


class OrcParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("damping", C.c_double), ("gravity", C.c_double * 3), ("verletDist", C.c_double),
        ("distFactor", C.c_double), ("poisson", C.c_double), ("en", C.c_double), ("gamma", C.c_double),
        ("alpha", C.c_double),
        ("law", C.c_int32), ("noSlip", C.c_int32), ("noBreak", C.c_int32), ("noFrict", C.c_int32),
        ("noAttraction", C.c_int32), ("trackEnergy", C.c_int32), ("trsfRenorm", C.c_int32), ("dist00", C.c_int32),
        ("loneMask", C.c_uint32), ("nThreads", C.c_int32),
    ]


_L = None


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _L
    if _L is not None:
        return _L
    if not os.path.exists(ORACLE_SO):
        build_oracle()
    L = C.CDLL(ORACLE_SO)
    P, I32, I64, D = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    pd, pi, pu, pl = C.POINTER(D), C.POINTER(I32), C.POINTER(C.c_uint32), C.POINTER(I64)
    sig = {
        "orc_create": (P, []), "orc_destroy": (None, [P]), "orc_last_error": (C.c_char_p, [P]),
        "orc_set_nodes": (C.c_int, [P, I32, pd, pd, pd, pd, pd, pd, pu, pd, pd]),
        "orc_set_particles": (C.c_int, [P, I32, pi, pi, pd, pi, pi, pu, pi]),
        "orc_set_materials": (C.c_int, [P, I32, pd, pd, pd, pd]),
        "orc_set_cell": (C.c_int, [P, I32, pd, pd, pd, I32]),
        "orc_set_params": (C.c_int, [P, C.POINTER(OrcParams)]),
        "orc_set_clumps": (C.c_int, [P, I32, pi, pi, pi, pd, pd]),
        "orc_step": (C.c_int, [P, I32]),
        "orc_get_nodes": (C.c_int, [P, pd, pd, pd, pd, pd, pd]),
        "orc_num_contacts": (I64, [P]),
        "orc_get_contacts": (C.c_int, [P, I64, pi, pi, pi, C.POINTER(C.c_uint8), pd, pd, pd, pd, pl]),
        "orc_get_energy": (C.c_int, [P, pd, pi]),
        "orc_get_stats": (C.c_int, [P, pl, pl, pd]),
        "orc_get_sort_stats": (C.c_int, [P, pl, pl]),
        "orc_get_cell": (C.c_int, [P, pd, pd]),
        "orc_bruteforce_pairs": (I64, [P, I64, pi, pi]),
        "orc_pwave_dt": (D, [P]),
        "orc_set_dyndt": (C.c_int, [P, C.c_void_p]),
        "orc_get_dyndt": (C.c_int, [P, C.c_void_p]),
        "orc_dyndt_crit": (C.c_int, [P, pd]),
        "orc_get_time": (C.c_int, [P, pd, pd]),
        "orc_remove_particles": (C.c_int, [P, I32, pi]),
        "orc_insert_spheres": (C.c_int, [P, I32, pd, pd, pd, pd, pd, pd, pi, pu, pi]),
        "orc_set_masks": (C.c_int, [P, I32, pi, pu]),
        "orc_get_free_ids": (C.c_int, [P, I32, pi, pi, pi]),
        "orc_make_cloud": (C.c_int, [pd, pd, D, D, I32, I32, D, pd, pd, I32, I32, pu, I64, pd, pd, pl, pd]),
        "orc_regular_pack": (C.c_int, [I32, I32, pd, D, D, I64, pd, pl]),
        "orc_set_contacts": (C.c_int, [P, I64, pi, pi, pi, pd, pd, pd, pd, pd]),
        "orc_set_time": (C.c_int, [P, I64, D]),
        "orc_get_contact_phys": (C.c_int, [P, I64, pd, pd]),
    }
    for n, (r, a) in sig.items():
        f = getattr(L, n)
        f.restype, f.argtypes = r, a
    _L = L
    return L


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def orc_params_from(wp, nThreads=0):
    """Copy a woodem_amd WdmParams (or any object with the same fields) into OrcParams."""
    p = OrcParams()
    for name, _ in OrcParams._fields_:
        if name == "nThreads":
            continue
        if name == "gravity":
            for k in range(3):
                p.gravity[k] = wp.gravity[k]
        else:
            setattr(p, name, getattr(wp, name))
    p.nThreads = nThreads
    return p


class OracleRun:
    """Same interface as woodem_amd.runner.HipRun, on the CPU oracle."""

    def __init__(self, arrays, params, cell, nThreads=0):
        L = lib()
        self.L = L
        self.h = L.orc_create()
        a = arrays.normalized()
        self.a = a
        D, I32, U32 = C.c_double, C.c_int32, C.c_uint32
        self._chk(L.orc_set_nodes(self.h, a.n_node, _p(a.pos, D), _p(a.ori, D), _p(a.vel, D), _p(a.angVel, D), _p(a.mass, D),
                                  _p(a.inertia, D), _p(a.flags, U32), _p(a.force, D), _p(a.torque, D)))
        self._chk(L.orc_set_particles(self.h, a.n_par, _p(a.shape, I32), _p(a.nodes, I32), _p(a.radius, D), _p(a.axis, I32),
                                      _p(a.sense, I32), _p(a.mask, U32), _p(a.mat, I32)))
        self._chk(L.orc_set_materials(self.h, a.n_mat, _p(a.density, D), _p(a.young, D), _p(a.ktDivKn, D), _p(a.tanPhi, D)))
        c = a.clumps
        if c:
            self._chk(L.orc_set_clumps(self.h, len(c["master"]), _p(c["master"], I32), _p(c["off"], I32), _p(c["members"], I32),
                                       _p(c["relPos"], D), _p(c["relOri"], D)))
        h9 = np.ascontiguousarray(cell["hSize"], np.float64).reshape(9)
        g9 = np.ascontiguousarray(cell["gradV"], np.float64).reshape(9)
        n9 = np.ascontiguousarray(cell["nextGradV"], np.float64).reshape(9)
        self._chk(L.orc_set_cell(self.h, int(cell["periodic"]), _p(h9, D), _p(g9, D), _p(n9, D), int(cell["homoDeform"])))
        op = params if isinstance(params, OrcParams) else orc_params_from(params, nThreads)
        self._chk(L.orc_set_params(self.h, C.byref(op)))

    def set_params(self, params, nThreads=0):
        op = params if isinstance(params, OrcParams) else orc_params_from(params, nThreads)
        self._chk(self.L.orc_set_params(self.h, C.byref(op)))

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError("oracle: " + self.L.orc_last_error(self.h).decode())

    def step(self, n=1):
        self._chk(self.L.orc_step(self.h, int(n)))

    def nodes(self):
        n = self.a.n_node
        out = {k: np.empty((n, 3)) for k in ("pos", "vel", "angVel", "force", "torque")}
        out["ori"] = np.empty((n, 4))
        D = C.c_double
        self._chk(self.L.orc_get_nodes(self.h, _p(out["pos"], D), _p(out["ori"], D), _p(out["vel"], D), _p(out["angVel"], D),
                                       _p(out["force"], D), _p(out["torque"], D)))
        return out

    def contacts(self):
        cap = int(self.L.orc_num_contacts(self.h))
        D, I32 = C.c_double, C.c_int32
        idA, idB = np.empty(cap, np.int32), np.empty(cap, np.int32)
        cd = np.empty((cap, 3), np.int32)
        real = np.empty(cap, np.uint8)
        fl, cp, tr, un = np.empty((cap, 3)), np.empty((cap, 3)), np.empty((cap, 9)), np.empty(cap)
        nout = C.c_int64()
        self._chk(self.L.orc_get_contacts(self.h, cap, _p(idA, I32), _p(idB, I32), _p(cd, I32),
                                          real.ctypes.data_as(C.POINTER(C.c_uint8)), _p(fl, D), _p(cp, D), _p(tr, D), _p(un, D),
                                          C.byref(nout)))
        return dict(idA=idA, idB=idB, cellDist=cd, isReal=real.astype(bool), forceLocal=fl, contPt=cp,
                    trsf=tr.reshape(-1, 3, 3), uN=un)

    def set_contacts(self, con):
        """real contacts of a checkpoint (woodem_amd.snapshot.read's contacts dict)"""
        D, I32 = C.c_double, C.c_int32
        g = lambda k, t: np.ascontiguousarray(con[k], t)
        idA, idB, cd = g("idA", np.int32), g("idB", np.int32), g("cellDist", np.int32)
        tr, fl, ph, ln, cp = g("trsf", np.float64), g("forceLocal", np.float64), g("phys", np.float64), g("lens", np.float64), \
            g("contPt", np.float64)
        self._chk(self.L.orc_set_contacts(self.h, len(idA), _p(idA, I32), _p(idB, I32), _p(cd, I32), _p(tr, D), _p(fl, D), _p(ph, D),
                                          _p(ln, D), _p(cp, D)))

    def contact_phys(self):
        """phys and lens of the contacts (order of contacts())"""
        cap = int(self.L.orc_num_contacts(self.h))
        ph, ln = np.empty((cap, 5)), np.empty((cap, 2))
        self._chk(self.L.orc_get_contact_phys(self.h, cap, _p(ph, C.c_double), _p(ln, C.c_double)))
        return ph, ln

    def set_time(self, step, time):
        self._chk(self.L.orc_set_time(self.h, int(step), float(time)))

    def energy(self):
        v = np.zeros(8)
        c = np.zeros(8, np.int32)
        self._chk(self.L.orc_get_energy(self.h, _p(v, C.c_double), _p(c, C.c_int32)))
        return v, c.astype(bool)

    def stats(self):
        s, f, vd = C.c_int64(), C.c_int64(), C.c_double()
        self._chk(self.L.orc_get_stats(self.h, C.byref(s), C.byref(f), C.byref(vd)))
        return dict(step=s.value, nFullRuns=f.value, verletDist=vd.value)

    def sort_stats(self):
        ch, pa = C.c_int64(), C.c_int64()
        self._chk(self.L.orc_get_sort_stats(self.h, C.byref(ch), C.byref(pa)))
        return dict(sortChunks=ch.value, sortPasses=pa.value)

    def cell(self):
        h, g = np.zeros(9), np.zeros(9)
        self._chk(self.L.orc_get_cell(self.h, _p(h, C.c_double), _p(g, C.c_double)))
        return h.reshape(3, 3), g.reshape(3, 3)

    def bruteforce_pairs(self):
        n = int(self.L.orc_bruteforce_pairs(self.h, 0, None, None))
        if n < 0:
            raise RuntimeError("bruteforce failed")
        a, b = np.empty(n, np.int32), np.empty(n, np.int32)
        self.L.orc_bruteforce_pairs(self.h, n, _p(a, C.c_int32), _p(b, C.c_int32))
        return a, b

    def set_dyndt(self, d):
        """DynDt engine (a woodem_amd._native.WdmDynDt: orc_dyndt has the same layout)."""
        self._chk(self.L.orc_set_dyndt(self.h, C.addressof(d)))

    def dyndt(self):
        from woodem_amd._native import WdmDynDt
        d = WdmDynDt()
        self._chk(self.L.orc_get_dyndt(self.h, C.addressof(d)))
        return d

    def dyndt_crit(self):
        v = C.c_double()
        self._chk(self.L.orc_dyndt_crit(self.h, C.byref(v)))
        return v.value

    def time(self):
        t, d = C.c_double(), C.c_double()
        self._chk(self.L.orc_get_time(self.h, C.byref(t), C.byref(d)))
        return t.value, d.value

    def remove_particles(self, ids):
        ids = np.ascontiguousarray(ids, np.int32)
        self._chk(self.L.orc_remove_particles(self.h, len(ids), _p(ids, C.c_int32)))

    def insert_spheres(self, pos, radius, mass, inertia, mat, mask, vel=None, angVel=None):
        f64 = lambda x: None if x is None else np.ascontiguousarray(x, np.float64)
        pos, vel, angVel = f64(pos).reshape(-1, 3), f64(vel), f64(angVel)
        n = len(pos)
        bc = lambda x, t: np.ascontiguousarray(np.broadcast_to(np.asarray(x, t), (n,)))
        radius, mass, inertia = bc(radius, np.float64), bc(mass, np.float64), bc(inertia, np.float64)
        mat, mask = bc(mat, np.int32), bc(mask, np.uint32)
        ids = np.empty(n, np.int32)
        D = C.c_double
        self._chk(self.L.orc_insert_spheres(self.h, n, _p(pos, D), _p(vel, D), _p(angVel, D), _p(radius, D), _p(mass, D),
                                            _p(inertia, D), _p(mat, C.c_int32), _p(mask, C.c_uint32), _p(ids, C.c_int32)))
        return ids

    def set_masks(self, ids, masks):
        ids = np.ascontiguousarray(ids, np.int32)
        masks = np.ascontiguousarray(np.broadcast_to(np.asarray(masks, np.uint32), ids.shape))
        self._chk(self.L.orc_set_masks(self.h, len(ids), _p(ids, C.c_int32), _p(masks, C.c_uint32)))

    def free_ids(self):
        n, e = C.c_int32(), C.c_int32()
        self._chk(self.L.orc_get_free_ids(self.h, 0, None, C.byref(n), C.byref(e)))
        ids = np.empty(n.value, np.int32)
        self._chk(self.L.orc_get_free_ids(self.h, n.value, _p(ids, C.c_int32), C.byref(n), C.byref(e)))
        return ids, e.value

    def pwave_dt(self):
        return float(self.L.orc_pwave_dt(self.h))

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pair_keys(c):
    """Canonical (min,max) pair keys of a contacts dict, as sorted int64."""
    lo = np.minimum(c["idA"], c["idB"]).astype(np.int64)
    hi = np.maximum(c["idA"], c["idB"]).astype(np.int64)
    return np.sort((lo << 32) | hi)


def make_cloud(mn, mx, rMean=-1, rRelFuzz=0, num=-1, periodic=False, porosity=.5, psdSizes=(), psdCumm=(), distributeMass=False,
               state=1, cap=100000):
    """oracle SpherePack::makeCloud: (centres, radii, final generator state, appliedPsdScaling)"""
    L = lib()
    f64 = lambda x: np.ascontiguousarray(x, np.float64)
    mn, mx, ps, pc = f64(mn), f64(mx), f64(psdSizes), f64(psdCumm)
    st = np.array([state], np.uint32)
    c = np.empty((cap, 3))
    r = np.empty(cap)
    n = C.c_int64()
    sc = C.c_double()
    D = C.c_double
    rc = L.orc_make_cloud(_p(mn, D), _p(mx, D), float(rMean), float(rRelFuzz), int(num), int(bool(periodic)), float(porosity),
                          _p(ps, D) if len(ps) else None, _p(pc, D) if len(pc) else None, len(ps), int(bool(distributeMass)),
                          _p(st, C.c_uint32), int(cap), _p(c, D), _p(r, D), C.byref(n), C.byref(sc))
    if rc != 0:
        raise RuntimeError("oracle makeCloud failed")
    return c[:n.value], r[:n.value], int(st[0]), sc.value


def regular_pack(hexa, kind, params, radius, gap, cap=1000000):
    """oracle pack.regularOrtho (hexa false) / regularHexa: the centres, in the reference's order"""
    L = lib()
    prm = np.zeros(7)
    prm[:len(params)] = params
    c = np.empty((cap, 3))
    n = C.c_int64()
    if L.orc_regular_pack(int(bool(hexa)), int(kind), _p(prm, C.c_double), float(radius), float(gap), int(cap), _p(c, C.c_double),
                          C.byref(n)) != 0:
        raise RuntimeError("oracle regular packing: %d points exceed cap" % n.value)
    return c[:n.value]
