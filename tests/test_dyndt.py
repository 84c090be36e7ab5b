"""DynDt (pkg/dem/DynDt.cpp:16-97) with the PeriodicEngine rule it runs by
(core/Engine.cpp:169-195).

The oracle's restatement is pinned by closed forms: the critical step of one contact along
an axis, with free and with blocked DOFs, for FrictPhys and HertzPhys stiffnesses; the
maxRelInc cap, dtSafety and dryRun of DynDt::run; the activation pattern of stepPeriod,
stepModulo, nDo and initRun; +inf (dt kept) without contacts.  The HIP path is compared
with the oracle: the same dt trajectory, bit for bit, with the windows of steps cut at the
steps DynDt fires on, and the same critical step on a dense scene.
"""
import math

import numpy as np
import pytest

import oracle_lib as OL
from woodem_amd import runner, scenes, soa
from woodem_amd._native import LAW_HERTZ_DMT, LAW_IDEALELPL, make_dyndt

R, RHO, E, KTDIVKN = 5e-3, 3000., 1e7, .2


def pair(overlap=2e-5, axis=0, flags=(0, 0), law=LAW_IDEALELPL, dt=1e-6):
    """two free spheres of radius R touching along `axis`, at rest"""
    pos = np.zeros((2, 3))
    pos[1, axis] = 2 * R - overlap
    a = soa.spheres_arrays(pos, R, density=RHO, young=E, ktDivKn=KTDIVKN, tanPhi=.5)
    a.flags[:] = flags
    p = runner.make_params(dt=dt, law=law, poisson=.3, en=float("nan"))
    return a, p, runner.cell_from_meta(a)


def frict_pair_expected(uN, flags=(0, 0)):
    """critical step of the pair after the creation step, from FrictPhys (FrictMat.cpp:21-44)
    and the DynDt stiffness of one contact along x: k_t = (kn, kt, kt), k_r = l^2 kt (0, 1, 1)"""
    m, inertia = soa.sphere_mass_inertia(R, RHO)
    l = R + uN / 2
    A = math.pi * R * R
    kn = 1 / (1 / (E * A / l) + 1 / (E * A / l))
    kt = KTDIVKN * kn
    cands = []
    for f in flags:
        if not f & 1:
            cands.append(m / kn)
        if not f & 2:
            cands.append(m / kt)
        if not f & 4:
            cands.append(m / kt)
        for ax in (1, 2):
            if not f & (8 << ax):
                cands.append(inertia / (l * l * kt))
    return math.sqrt(2 * min(cands)), kn, kt


def test_crit_frict_pair_closed_form():
    a, p, cell = pair()
    o = OL.OracleRun(a, p, cell)
    assert o.dyndt_crit() == math.inf  # no contacts yet
    o.step(1)
    c = o.contacts()
    assert c["isReal"].sum() == 1
    want, kn, kt = frict_pair_expected(float(c["uN"][0]))
    assert o.dyndt_crit() == pytest.approx(want, rel=1e-14)
    # the translational bound wins over the rotational one here
    m, inertia = soa.sphere_mass_inertia(R, RHO)
    assert want == pytest.approx(math.sqrt(2 * m / kn), rel=1e-14)


def test_crit_blocked_dofs_closed_form():
    # x translation blocked on both: the next bound is rotation about y/z or translation in y/z
    a, p, cell = pair(flags=(1, 1))
    o = OL.OracleRun(a, p, cell)
    o.step(1)
    uN = float(o.contacts()["uN"][0])
    want, kn, kt = frict_pair_expected(uN, flags=(1, 1))
    assert o.dyndt_crit() == pytest.approx(want, rel=1e-14)
    # every DOF blocked: infinite
    a, p, cell = pair(flags=(63, 63))
    o = OL.OracleRun(a, p, cell)
    o.step(1)
    assert o.dyndt_crit() == math.inf


def test_crit_hertz_pair_closed_form():
    a, p, cell = pair(law=LAW_HERTZ_DMT)
    o = OL.OracleRun(a, p, cell)
    o.step(1)
    uN = float(o.contacts()["uN"][0])
    nu = .3
    # Cp2_FrictMat_HertzPhys::go (Hertz.cpp:9-50) on lens l = R + uN0/2 (uN0 at creation = uN)
    l = R + uN / 2
    K = 4 / 3. / (2 * (1 - nu * nu) / E)
    Rr = l / 2
    G = E / (2 * (1 + nu))
    kt0 = 2 * math.sqrt(4 * Rr) * G / (2 - nu)
    kn = 1.5 * K * math.sqrt(Rr) * math.sqrt(-uN)
    kt = kt0 * math.sqrt(-uN)
    m, inertia = soa.sphere_mass_inertia(R, RHO)
    want = math.sqrt(2 * min(m / kn, m / kt, inertia / (l * l * kt)))
    assert o.dyndt_crit() == pytest.approx(want, rel=1e-12)


def settled_box(n=(6, 5, 4)):
    a = scenes.box_lattice(nx=n[0], ny=n[1], nz=n[2])
    return a, runner.params_from_meta(a), runner.cell_from_meta(a)


def settled_oracle():
    """the box after 2500 steps: about 50 real contacts (spheres on the floor and each other)"""
    a, p, cell = settled_box()
    o = OL.OracleRun(a, p, cell)
    o.step(2500)
    assert o.contacts()["isReal"].sum() > 20
    return o, p


def test_maxrelinc_cap():
    o, p = settled_oracle()
    crit = o.dyndt_crit()
    assert math.isfinite(crit)
    # dt far below the critical one: a run may raise it by (1+maxRelInc)^(steps since the last run)
    dt0 = crit * 1e-3
    p.dt = dt0
    o.set_params(p)
    s0 = o.stats()["step"]
    assert s0 % 10 == 0
    o.set_dyndt(make_dyndt(stepPeriod=10, maxRelInc=1e-2, stepLast=s0 - 10))
    o.step(1)  # fires at step s0: nSteps = 10 since stepPrev
    assert o.time()[1] == dt0 * math.pow(1.01, 10)
    d = o.dyndt()
    assert (d.nDone, d.stepLast, d.stepPrev) == (1, s0, s0 - 10)


def test_dtsafety_bound():
    # dt above crit * dtSafety: set to it at the next run (first run: nSteps = 0, no growth)
    o, p = settled_oracle()
    o.set_dyndt(make_dyndt(stepPeriod=1, dtSafety=.1))
    dt = o.time()[1]
    o.step(1)
    crit = o.dyndt_crit()  # the run saw this state: contacts after the step's ContactLoop
    assert crit * .1 < dt
    assert o.time()[1] == crit * .1


def test_dryrun_keeps_dt():
    o, p = settled_oracle()
    dt = o.time()[1]
    o.set_dyndt(make_dyndt(stepPeriod=1, dryRun=1, dtSafety=.1))
    o.step(1)
    crit = o.dyndt_crit()
    assert o.time()[1] == dt
    assert o.dyndt().dt == crit * .1


def test_periodic_engine_activation():
    a, p, cell = pair()
    # stepPeriod 7 with stepModulo: fires when step % 7 == 0 (the first call only starts counting)
    o = OL.OracleRun(a, p, cell)
    o.step(3)
    o.set_dyndt(make_dyndt(stepPeriod=7, maxRelInc=1e-9))
    fired = []
    for _ in range(30):
        n0 = o.dyndt().nDone
        s = o.stats()["step"]
        o.step(1)
        if o.dyndt().nDone > n0:
            fired.append(s)
    assert fired == [7, 14, 21, 28]
    # nDo limits the runs; without stepModulo the period counts from the last run
    o = OL.OracleRun(a, p, cell)
    o.step(3)
    o.set_dyndt(make_dyndt(stepPeriod=7, stepModulo=0, nDo=2))
    fired = []
    for _ in range(40):
        n0 = o.dyndt().nDone
        s = o.stats()["step"]
        o.step(1)
        if o.dyndt().nDone > n0:
            fired.append(s)
    assert fired == [10, 17]
    # initRun=False: the first activation only records the counters
    o = OL.OracleRun(a, p, cell)
    o.set_dyndt(make_dyndt(stepPeriod=5, initRun=0))
    o.step(1)  # step 0: 0 % 5 == 0 but it is the initialising call
    d = o.dyndt()
    assert (d.nDone, d.stepLast) == (0, 0)
    o.step(5)
    assert o.dyndt().nDone == 1 and o.dyndt().stepLast == 5


def test_no_contacts_keeps_dt():
    a, p, cell = pair(overlap=-1e-3)  # apart
    o = OL.OracleRun(a, p, cell)
    o.set_dyndt(make_dyndt(stepPeriod=1))
    o.step(5)
    assert o.dyndt().nDone == 5
    assert o.time()[1] == p.dt


# ---------------- HIP path against the oracle ----------------
@pytest.mark.gpu
@pytest.mark.parametrize("window", [1, 16])
def test_gpu_dt_trajectory_equals_oracle(window):
    from woodem_amd._native import Context
    ctx = Context(0)
    ctx.set_tuning(window=window)
    a, p, cell = settled_box()
    p.trackEnergy = 1
    g = runner.HipRun(a, p, cell, ctx=ctx)
    o = OL.OracleRun(a, p, cell)
    dd = make_dyndt(stepPeriod=9, maxRelInc=2e-3, dtSafety=.6)
    g.ctx.set_dyndt(g.sid, dd)
    o.set_dyndt(dd)
    for n in (1, 40, 137, 300, 2000):
        g.step(n)
        o.step(n)
        tg, dg = g.ctx.time(g.sid)
        to, do = o.time()
        assert dg == do and tg == to
        gd, od = g.ctx.dyndt(g.sid), o.dyndt()
        assert (gd.nDone, gd.stepLast, gd.stepPrev) == (od.nDone, od.stepLast, od.stepPrev)
    assert g.ctx.dyndt(g.sid).nDone > 40
    ng, no = g.nodes(), o.nodes()
    for q in ("pos", "vel", "angVel"):
        assert np.array_equal(ng[q], no[q]), q
    assert g.ctx.dyndt_crit(g.sid) == o.dyndt_crit()
    g.close()


@pytest.mark.gpu
def test_gpu_crit_equals_oracle_periodic_hertz():
    a = scenes.periodic_hertz(n=3000, seed=5)
    from woodem_amd._native import Context
    ctx = Context(0)
    b = scenes.relax_and_order(a, ctx, iters=100)
    p = runner.params_from_meta(b)
    cell = runner.cell_from_meta(b)
    g = runner.HipRun(b, p, cell, ctx=ctx)
    o = OL.OracleRun(b, p, cell)
    dd = make_dyndt(stepPeriod=25)
    g.ctx.set_dyndt(g.sid, dd)
    o.set_dyndt(dd)
    g.step(120)
    o.step(120)
    cg, co = g.ctx.dyndt_crit(g.sid), o.dyndt_crit()
    assert math.isfinite(co) and cg == co
    assert g.ctx.time(g.sid) == o.time()
    g.close()


@pytest.mark.gpu
def test_gpu_scene_api_default_engines():
    """utils.defaultEngines includes DynDt(stepPeriod=100) (py/utils.py:83): the Scene's dt and
    the engine's counters follow the oracle's"""
    import woodem_amd as woo
    from woodem_amd import utils
    S = woo.Scene(fields=[woo.DemField(gravity=(0, 0, -9.81))], dt=2.6e-5)
    S.dem.par.add(utils.wallBox(((0, 0, 0), (.1, .1, .2))))
    S.dem.par.add([utils.sphere((.05 + .001 * (k % 3), .05, .006 + .0105 * k), .005) for k in range(12)])
    S.engines = utils.defaultEngines(damping=.2)
    a = S.dem._compile()
    p = S._params()
    dd = S.engines[-1]._native(S.dtSafety)
    S.run(1700)
    o = OL.OracleRun(a, p, runner.cell_from_meta(a))
    o.set_dyndt(dd)
    o.step(1700)
    assert S.engines[-1].nDone == o.dyndt().nDone == 17
    assert (S.time, S.dt) == o.time()
    assert S.dt != 2.6e-5  # the spheres touch: DynDt changed dt
    no = o.nodes()
    for i in range(a.n_par):
        assert np.array_equal(S.dem.par[i].pos, no["pos"][i])
