"""Closed-form pins of the oracle's contact-law energies and of the rotation update.

Kinematic contacts: both bodies have every DOF blocked, so Leapfrog moves them at constant
velocity (Leapfrog.cpp:275-278 still integrates positions) and the contact sees an exactly
known overlap history.  A sphere against a wall keeps the contact normal at e_z, so the
L6Geom local velocity is exactly (v_n, -v_x, -v_y) (the frame (e_z, -e_x, -e_y) of
L6Geom::setInitialLocalCoords, L6Geom.cpp:27-38, with no rotation).  The laws' energies then
follow from the mechanics of their springs and dashpots:

* IdealElPl (IdealElPl.cpp:40-77): once the shear spring slides, every step dissipates the
  friction force times the slip distance, maxFt v_t dt, in `plast`; `elast` is the energy of
  the two linear springs, (Fn^2/kn + Ft^2/kt)/2.
* Hertz/DMT (Hertz.cpp:58-182): `elast` is the Hertz potential 2/5 kn0 delta^(5/2) (the
  integral of kn0 delta^(3/2)) plus the shear spring's Ft^2/(2 kt); `viscN` and `viscT` are
  the work of the normal and tangential dashpots, eta v^2 dt with eta = alpha_sqrtMK
  delta^(1/4) (Hertz.cpp:36-44); `plast` is the shear-spring energy released by the Coulomb
  cap, (|Ft_trial|^2 - maxFt^2) / (2 kt).  The Hertz potential is also checked against the
  work of the elastic force along the path (sum of Fne d(delta)), which no rewriting of the
  law's own expression can satisfy with a wrong coefficient.
* A free collision loses exactly what the dashpot dissipates: initial KE = final KE + viscN.
* Leapfrog rotates a node by q <- AngleAxis(|w| dt, w/|w|) q (Leapfrog.cpp:280-288): spins
  about two axes in sequence compose as dq2^N2 dq1^N1 q0, and the reversed product differs.

Plausible slips that fail here: 2/3 for 2/5 in elast, the commented-out
(.5(Ft-Fm)+Fm)(Ft-Fm)/kt for IdealElPl's plast (IdealElPl.cpp:69), a dropped dashpot term,
a wrong viscosity exponent, q dq for dq q.
"""
import math

import numpy as np
import pytest

import oracle_lib as OL
from woodem_amd import runner, soa
from woodem_amd._native import LAW_HERTZ_DMT

NOCELL = dict(periodic=False, hSize=np.eye(3), gradV=np.zeros((3, 3)), nextGradV=np.zeros((3, 3)), homoDeform=4)
E_ = {n: i for i, n in enumerate(OL.ENERGY_NAMES)}


def kinematic_sphere_on_floor(r, z0, vel, density=1e3, young=1e7, tanPhi=.5, ktDivKn=.2):
    """one sphere (every DOF blocked, constant velocity) at height z0 over the wall z = 0"""
    a = soa.spheres_arrays(np.array([[.1, .1, z0]]), r, density=density, young=young, tanPhi=tanPhi, ktDivKn=ktDivKn,
                           vel=np.array([vel], float))
    a = soa.append_walls(a, [(0., 2, 1)])
    a.flags[0] = soa.DOF_ALL
    return a


def test_idealelpl_plastic_dissipation_is_friction_work():
    r, E, delta, vt, dt = .01, 1e7, 1e-4, .1, 1e-6
    a = kinematic_sphere_on_floor(r, r - delta, (vt, 0., 0.), young=E)
    p = runner.make_params(dt=dt, damping=0., trackEnergy=True)
    o = OL.OracleRun(a, p, NOCELL)
    S = 6000
    o.step(S)
    c = o.contacts()
    assert c["isReal"].all() and c["uN"][0] == (r - delta) - r
    uN = c["uN"][0]
    # FrictPhys of a wall-sphere contact (FrictMat.cpp:21-44 with L6Geom lens = r + uN/2 on both
    # sides, contA = pi r^2): the two springs in series
    l = r + uN / 2
    kn = E * math.pi * r * r / (2 * l)
    kt = .2 * kn
    Fn = kn * uN
    maxFt = abs(Fn) * .5
    F = c["forceLocal"][0]
    assert F[0] == pytest.approx(Fn, rel=1e-12)
    assert math.hypot(F[1], F[2]) == pytest.approx(maxFt, rel=1e-12)  # sliding: on the Coulomb cone
    # local y is -e_x and the sphere (side B) receives -trsf F: -|Ft| e_x opposes the slide
    assert F[1] < 0 and F[2] == 0
    # the shear spring loads by d = dt kt vt per step from the step after creation (k = 1..S-1)
    # until it passes maxFt at k* (the excess d k* - maxFt is released), then it slides
    d = dt * kt * vt
    kstar = int(math.floor(maxFt / d)) + 1
    assert 100 < kstar < S - 1000
    plast = maxFt * (kstar * d - maxFt) / kt + (S - 1 - kstar) * maxFt * vt * dt
    v, cr = o.energy()
    assert cr[E_["plast"]] and v[E_["plast"]] == pytest.approx(plast, rel=1e-9)
    # the alternative (.5(Ft-Fm)+Fm)(Ft-Fm)/kt adds kt (vt dt)^2/2 per sliding step: far outside
    alt = plast + (S - 1 - kstar) * .5 * kt * (vt * dt) ** 2
    assert abs(alt - plast) > 1e-5 * plast
    assert v[E_["elast"]] == pytest.approx(.5 * (Fn * Fn / kn + maxFt * maxFt / kt), rel=1e-12)


def _hertz_constants(r, uN0, E, nu, en, m):
    """HertzPhys of a wall-sphere contact created at overlap -uN0 (Hertz.cpp:9-50): lens
    r + uN0/2 on both sides, the wall massless (the sphere's mass is the effective one)."""
    l = r + uN0 / 2
    R = l / 2
    K = 4. / 3. * E / (2 * (1 - nu * nu))
    kn0 = K * math.sqrt(R)
    G = E / (2 * (1 + nu))
    kt0 = 2 * math.sqrt(4 * R) * G / (2 - nu)
    le = math.log(en)
    alpha = -math.sqrt(5) * le / math.sqrt(le * le + math.pi ** 2)
    return kn0, kt0, alpha * math.sqrt(m * kn0)


@pytest.mark.parametrize("vn,vt,dt", [(.05, 2e-3, 2e-8), (2e-3, .2, 1e-7)])
def test_hertz_energies_closed_form(vn, vt, dt):
    """A sphere leaving the floor at v_n while sliding at v_t.  Fast separation (v_n .05): the
    viscous normal force exceeds the elastic one (noAttraction off), Fn > 0 and the shear stays
    under the cap, so viscT accumulates; slow separation (v_n 2e-3) with a fast slide: Fn < 0
    until the last tenth of the path, the cap is 0 and every step releases the trial shear
    into plast."""
    r, rho, E, nu, en, tanPhi = .01, 2650., 7e10, .3, .2, .4
    delta0 = 2e-6
    a = kinematic_sphere_on_floor(r, r - delta0, (vt, 0., vn), density=rho, young=E, tanPhi=tanPhi)
    m = a.mass[0]
    p = runner.make_params(dt=dt, damping=0., law=LAW_HERTZ_DMT, poisson=nu, en=en, noAttraction=False, trackEnergy=True)
    o = OL.OracleRun(a, p, NOCELL)
    # positions as Leapfrog advances them (z += vn dt each step), and uN = z - r (Wall.cpp:59-64)
    z = r - delta0
    kn0 = kt0 = asq = None
    tot = dict(viscN=0., viscT=0., plast=0.)
    elast = 0.
    work = 0.       # work of the elastic normal force along the path, sum of Fne d(delta)
    pot0 = None
    nslip = nvisc = 0
    step = 0
    while True:
        z = z + vn * dt
        uN = z - r
        if uN > 0:
            break
        o.step(1)
        c = o.contacts()
        assert c["isReal"].all() and c["uN"][0] == uN
        if step == 0:
            kn0, kt0, asq = _hertz_constants(r, uN, E, nu, en, m)
            gv = np.zeros(3)  # a new contact has no relative velocity yet
        else:
            gv = np.array([vn, -vt, 0.])
        dl = -uN
        Fne = -kn0 * dl ** 1.5
        eta = asq * dl ** .25
        Fnv = eta * gv[0]
        Fn = Fne + Fnv
        tot["viscN"] += Fnv * gv[0] * dt
        kt = kt0 * math.sqrt(dl)
        Ft = dt * kt * gv[1:]
        maxFt = max(0., Fn) * tanPhi
        if Ft @ Ft > maxFt * maxFt:
            tot["plast"] += .5 * (Ft @ Ft - maxFt * maxFt) / kt
            Ft = Ft * (maxFt / math.sqrt(Ft @ Ft))
            nslip += 1
        else:
            Ft = Ft + eta * gv[1:]
            tot["viscT"] += eta * (gv[1:] @ gv[1:]) * dt
            nvisc += step > 0
        pot = .4 * kn0 * dl ** 2.5
        elast = pot + .5 * (Ft @ Ft) / kt
        if pot0 is None:
            pot0 = pot
        else:
            work += -Fne * (prev_dl - dl)
        prev_dl = dl
        F = c["forceLocal"][0]
        assert F[0] == pytest.approx(Fn, rel=1e-11, abs=1e-12 * abs(Fne))
        assert F[1:] == pytest.approx(Ft, rel=1e-11, abs=1e-14)
        v, cr = o.energy()
        assert v[E_["elast"]] == pytest.approx(elast, rel=1e-11)
        for k in ("viscN", "viscT", "plast"):
            assert v[E_[k]] == pytest.approx(tot[k], rel=1e-11, abs=1e-25), k
        step += 1
    assert step > 1000
    if vn > .01:
        assert nvisc > 1000 and tot["viscT"] > 0  # the branch under the cap was exercised
    else:
        assert nslip > .8 * step and tot["plast"] > 0
    # the Hertz potential released equals the elastic force's work along the path (trapezoid
    # error of the discrete sum ~ one step's share)
    assert pot0 - .4 * kn0 * prev_dl ** 2.5 == pytest.approx(work, rel=2e-3)


def test_hertz_collision_energy_balance():
    """Free head-on collision of two spheres with en < 1: what the normal dashpot dissipates is
    the kinetic energy lost (no friction, no plasticity in a head-on impact); at the deepest
    compression the Hertz potential holds what is neither kinetic nor dissipated.  The leapfrog
    step leaves a discretisation error of order dt/tau in the dissipated energy."""
    r, rho, E, nu, en = .01, 2650., 7e10, .3, .5
    v0 = .5
    a = soa.spheres_arrays(np.array([[0., 0., 0.], [2 * r + 1e-9, 0., 0.]]), r, density=rho, young=E,
                           vel=np.array([[v0 / 2, 0, 0], [-v0 / 2, 0, 0]]))
    m = a.mass[0]
    ke0 = 2 * .5 * m * (v0 / 2) ** 2
    tau = 2.214 * (rho / (E / (2 * (1 - nu * nu)))) ** .4 * (2 * r) * v0 ** -.2  # py/tests/hertz.py:38-46
    dt = tau / 4000
    p = runner.make_params(dt=dt, damping=0., law=LAW_HERTZ_DMT, poisson=nu, en=en, trackEnergy=True)
    o = OL.OracleRun(a, p, NOCELL)
    seen, best = False, None
    for _ in range(20000):
        o.step(1)
        c = o.contacts()
        v, cr = o.energy()
        if c["isReal"].any():
            seen = True
            if best is None or c["uN"][0] < best[0]:
                best = (c["uN"][0], v.copy())
        elif seen:
            break
    v, cr = o.energy()
    nd = o.nodes()
    ke1 = .5 * m * (nd["vel"] ** 2).sum()
    # kinetic slot and the nodes agree (Leapfrog's on-step velocity, no force after separation)
    assert v[E_["kinetic"]] == pytest.approx(ke1, rel=1e-12)
    assert .1 < ke1 / ke0 < .5  # about en^2 (the noAttraction clamp lets a little more rebound)
    assert cr[E_["viscN"]] and not cr[E_["plast"]] and not cr[E_["viscT"]]
    assert ke0 == pytest.approx(ke1 + v[E_["viscN"]], rel=2e-3)
    uNmax, vm = best
    assert vm[E_["elast"]] > .2 * ke0
    assert ke0 == pytest.approx(vm[E_["kinetic"]] + vm[E_["elast"]] + vm[E_["viscN"]], rel=2e-3)


def _qmul(a, b):
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def _aa(axis, angle):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    return np.concatenate([[math.cos(angle / 2)], math.sin(angle / 2) * axis])


def spin_in_sequence(run_factory, w1, w2, n1, n2, dt):
    """orientation after n1 steps at spin w1 then n2 steps at w2 (a free sphere, no gravity)"""
    a = soa.spheres_arrays(np.array([[0., 0., 0.], [1., 1., 1.]]), .01)
    a.angVel[0] = w1
    q0 = _aa([1, 2, 3], .7)
    a.ori[0] = q0
    r1 = run_factory(a, dt)
    r1.step(n1)
    nd = r1.nodes()
    b = a.copy()
    b.pos, b.ori, b.vel = nd["pos"], nd["ori"], nd["vel"]
    b.angVel = nd["angVel"].copy()
    b.angVel[0] = w2
    r2 = run_factory(b, dt)
    r2.step(n2)
    return r2.nodes()["ori"][0], q0


def test_rotation_composes_left():
    w1, w2, dt, n = np.array([0., 0., 3.]), np.array([2., 0., 0.]), 1e-4, 2000
    oracle = lambda a, dt: OL.OracleRun(a, runner.make_params(dt=dt), NOCELL)
    q, q0 = spin_in_sequence(oracle, w1, w2, n, n, dt)
    dq1 = _aa(w1, np.linalg.norm(w1) * dt * n)
    dq2 = _aa(w2, np.linalg.norm(w2) * dt * n)
    want = _qmul(dq2, _qmul(dq1, q0))
    twisted = _qmul(_qmul(q0, dq1), dq2)
    assert np.abs(want - twisted).max() > .1
    assert np.abs(q - want).max() < 1e-12


# ---------------- the HIP path on the pinned scenes ----------------
@pytest.mark.gpu
@pytest.mark.parametrize("case", ["idealelpl_slide", "hertz_visc", "hertz_slip", "spin"])
def test_gpu_matches_pinned_scenes(hip_ctx, case):
    """The same kinematic scenes through the C ABI: forces, energies and orientations equal to
    the oracle's, whose values the closed forms above pin."""
    if case == "idealelpl_slide":
        a = kinematic_sphere_on_floor(.01, .01 - 1e-4, (.1, 0., 0.))
        p, steps = runner.make_params(dt=1e-6, damping=0., trackEnergy=True), 6000
    elif case == "spin":
        a = soa.spheres_arrays(np.array([[0., 0., 0.], [1., 1., 1.]]), .01)
        a.angVel[0] = (0., 0., 3.)
        a.angVel[1] = (2., 1., -1.)
        a.ori[0] = _aa([1, 2, 3], .7)
        p, steps = runner.make_params(dt=1e-4), 2000
    else:
        vn, vt, dt = (.05, 2e-3, 2e-8) if case == "hertz_visc" else (2e-3, .2, 1e-7)
        a = kinematic_sphere_on_floor(.01, .01 - 2e-6, (vt, 0., vn), density=2650., young=7e10, tanPhi=.4)
        p = runner.make_params(dt=dt, damping=0., law=LAW_HERTZ_DMT, poisson=.3, en=.2, noAttraction=False, trackEnergy=True)
        steps = int(2e-6 / (vn * dt)) - 2
    g = runner.HipRun(a, p, NOCELL, ctx=hip_ctx)
    o = OL.OracleRun(a, p, NOCELL)
    for n in (1, steps // 2, steps - 1 - steps // 2):
        g.step(n)
        o.step(n)
        ng, no = g.nodes(), o.nodes()
        for q in ("pos", "ori", "vel", "angVel", "force", "torque"):
            assert np.array_equal(ng[q], no[q]), q
        vg, cg = g.energy()
        vo, co = o.energy()
        assert np.array_equal(cg, co) and np.array_equal(vg, vo)
    g.close()
