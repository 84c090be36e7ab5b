"""Work-list relayout between collider full runs (wdm_set_relayout): contacts that became real
after the full run moved right after its state rows in the ContactLoop's work list, with their
state rows, contact points and gather records, CSR entries remapped.  Memory order only: stepped with a relayout every 8 steps (2 in the
drum, whose full runs come every few steps), every scene
kind gives the same nodes, contact list and contact state bit for bit as without relayouts
(energies sum their terms in work-list order: equal to rounding), and stepping engine by engine
relayouts at the same steps as wdm_step, energies included; with particles removed and
inserted between relayouts the run still equals the oracle bit for bit."""
import numpy as np
import pytest

import oracle_lib as OL
from woodem_amd import runner, scenes, soa
from woodem_amd._native import PREC_MIXED, make_dyndt

pytestmark = pytest.mark.gpu


def _scene(kind, ctx):
    if kind == "box":
        a = scenes.relax_box(scenes.deposition_box(seed=1003), ctx, iters=200)
        return a, runner.params_from_meta(a, trackEnergy=True)
    if kind == "drum":  # spheres thrown at the spinning wall: facet, wall and sphere contacts early
        a = scenes.drum(seed=3, n=2000, omega=3., gap=1e-4)
        ns = a.meta["n_spheres"]
        rng = np.random.default_rng(5)
        a.vel[:ns] = rng.normal(0, .3, (ns, 3))
        a.vel[:ns, 2] -= .8
        return a, runner.params_from_meta(a, trackEnergy=True)
    a = scenes.relax_and_order(scenes.periodic_hertz(n=4000, seed=6), ctx, iters=150)
    if kind == "mixed":
        return a, runner.params_from_meta(a, trackEnergy=True, precision=PREC_MIXED)
    return a, runner.params_from_meta(a, trackEnergy=True)


def _run(ctx, a, p, steps, relayout, dyndt=False):
    ctx.set_relayout(relayout, 0)
    r = runner.HipRun(a, p, runner.cell_from_meta(a), ctx=ctx)
    if dyndt:
        ctx.set_dyndt(r.sid, make_dyndt(stepPeriod=50, maxRelInc=1e-3))
    r.step(steps)
    out = dict(nodes=r.nodes(), contacts=r.contacts(), energy=r.energy(), stats=r.stats(), relayouts=ctx.relayout(r.sid)[2],
               time=ctx.time(r.sid))
    r.close()
    return out


@pytest.mark.parametrize("kind", ["periodic", "box", "drum", "mixed"])
def test_relayout_changes_no_result(hip_ctx, kind):
    ctx = hip_ctx
    a, p = _scene(kind, ctx)
    steps = 1500 if kind == "drum" else 400
    try:
        # the drum's full runs come every few steps: relayouts every 2 steps fall between them
        x = _run(ctx, a, p, steps, 2 if kind == "drum" else 8, dyndt=kind == "box")
        y = _run(ctx, a, p, steps, 0, dyndt=kind == "box")
    finally:
        ctx.set_relayout(32)
    assert x["relayouts"] > 10 and y["relayouts"] == 0
    for q in ("pos", "ori", "vel", "angVel", "force", "torque"):
        assert np.array_equal(x["nodes"][q], y["nodes"][q]), q
    for q in ("idA", "idB", "cellDist", "isReal", "forceLocal", "contPt", "trsf", "uN"):
        assert np.array_equal(x["contacts"][q], y["contacts"][q], equal_nan=True), q
    assert x["contacts"]["isReal"].sum() > (20 if kind == "drum" else 100)
    assert x["stats"] == y["stats"] and x["time"] == y["time"]
    (ex, cx), (ey, cy) = x["energy"], y["energy"]
    assert np.array_equal(cx, cy)
    assert np.allclose(ex, ey, rtol=1e-11, atol=1e-11 * np.abs(ey).max())


def test_engine_by_engine_relayouts_like_wdm_step(hip_ctx):
    ctx = hip_ctx
    a, p = _scene("periodic", ctx)
    cell = runner.cell_from_meta(a)
    ctx.set_relayout(8, 0)
    try:
        eng = runner.HipRun(a, p, cell, ctx=ctx)
        whole = runner.HipRun(a, p, cell, ctx=ctx)
        for s in range(150):
            sid = eng.sid
            ctx.step_begin(sid)
            ctx.leapfrog_run(sid)
            ctx.collider_run(sid)
            ctx.contact_loop_run(sid)
            ctx.step_end(sid)
        whole.step(150)
        n1, n2 = eng.nodes(), whole.nodes()
        for q in ("pos", "vel", "angVel", "force", "torque"):
            assert np.array_equal(n1[q], n2[q]), q
        e1, e2 = eng.energy(), whole.energy()
        assert np.array_equal(e1[0], e2[0]) and np.array_equal(e1[1], e2[1])
        assert ctx.relayout(eng.sid)[2] == ctx.relayout(whole.sid)[2] > 5
        eng.close()
        whole.close()
    finally:
        ctx.set_relayout(32)


def test_relayouts_with_removal_and_insertion_equal_the_oracle(hip_ctx):
    """Relayouts every 2 steps around DemField::removeParticle / particle insertion
    (pkg/dem/Particle.cpp:445-493): positions and contact sets equal the oracle's bit for bit."""
    ctx = hip_ctx
    r, rho = 5e-3, 3000.
    ctx.set_relayout(2, 0)
    try:
        a = soa.with_spare(scenes.box_lattice(nx=8, ny=6, nz=5), 40)
        p = runner.params_from_meta(a, trackEnergy=True)
        cell = runner.cell_from_meta(a)
        g = runner.HipRun(a, p, cell, ctx=ctx)
        o = OL.OracleRun(a, p, cell)
        g.step(1500)
        o.step(1500)
        z = o.nodes()["pos"][:, 2]
        sph = np.nonzero(a.shape == soa.SHAPE_SPHERE)[0]
        rm = sph[np.argsort(-z[sph], kind="stable")[:25]]
        ctx.remove_particles(g.sid, rm)
        o.remove_particles(rm)
        g.step(200)
        o.step(200)
        m, inertia = soa.sphere_mass_inertia(r, rho)
        newpos = np.array([[.05 + .012 * k, .12, .2] for k in range(30)])
        assert np.array_equal(ctx.insert_spheres(g.sid, newpos, r, m, inertia, 0, soa.DEFAULT_MOVABLE_MASK),
                              o.insert_spheres(newpos, r, m, inertia, 0, soa.DEFAULT_MOVABLE_MASK))
        for n in (1, 50, 600):
            g.step(n)
            o.step(n)
            ng, no = g.nodes(), o.nodes()
            for q in ("pos", "vel", "angVel", "force", "torque"):
                assert np.array_equal(ng[q], no[q]), (n, q)
            assert np.array_equal(OL.pair_keys(g.contacts()), OL.pair_keys(o.contacts()))
        assert ctx.relayout(g.sid)[2] > 100
        g.close()
        o.close()
    finally:
        ctx.set_relayout(32)


def test_small_lists_keep_their_layout_and_windows(hip_ctx):
    """With the default settings (relayouts every 32 steps for lists of 2^20 contacts or more)
    a small scene is never relaid out and its windows are not cut short: 160 steps take 10
    windows of 16."""
    ctx = hip_ctx
    ctx.set_relayout(32)
    a, p = _scene("periodic", ctx)
    r = runner.HipRun(a, p, runner.cell_from_meta(a), ctx=ctx)
    r.step(40)  # the first full runs of the relaxed packing
    full0 = r.stats()["nFullRuns"]
    w0 = ctx.tuning()["windows"]
    r.step(160)
    if r.stats()["nFullRuns"] == full0:  # a Verlet trigger ends a window early
        assert ctx.tuning()["windows"] - w0 == 10
    assert ctx.relayout(r.sid)[1] == 1 << 20 and ctx.relayout(r.sid)[2] == 0
    r.close()
