"""The drop-in boundary engine by engine: a host Scene::doOneStep (core/Scene.cpp:319-424)
calling the C ABI once per engine, as the C++ engine classes of the integration do
(Engine::run, core/Engine.hpp:27-105): wdm_step_begin, wdm_leapfrog_run
(Leapfrog::run, pkg/dem/Leapfrog.cpp:103-255), wdm_collider_run
(InsertionSortCollider::run, pkg/dem/InsertionSortCollider.cpp:445-646),
wdm_contact_loop_run (ContactLoop::run, pkg/dem/ContactLoop.cpp:93-259), wdm_dyndt_run
(DynDt), wdm_step_end.  With read-backs between the calls, the trajectory equals the one of
wdm_step (whole steps, windows of graphs) bit for bit and the oracle's; between Leapfrog and
ContactLoop the node forces read back are Leapfrog's reset (m g, no torque,
Leapfrog.cpp:240-245), as DemData.force holds them in the reference."""
import numpy as np
import pytest

import oracle_lib as OL
from woodem_amd import runner, scenes, soa
from woodem_amd._native import make_dyndt

pytestmark = pytest.mark.gpu


def _scene(kind, ctx):
    if kind == "box":
        a = scenes.box_lattice(nx=8, ny=6, nz=5)
    else:
        a = scenes.relax_and_order(scenes.periodic_hertz(n=3000, seed=2), ctx, iters=150)
    return a, runner.params_from_meta(a, trackEnergy=True), runner.cell_from_meta(a)


def _same(r1, r2, energy=True):
    n1, n2 = r1.nodes(), r2.nodes()
    for q in ("pos", "ori", "vel", "angVel", "force", "torque"):
        assert np.array_equal(n1[q], n2[q]), q
    c1, c2 = r1.contacts(), r2.contacts()
    for k in ("idA", "idB", "cellDist", "isReal", "forceLocal"):
        assert np.array_equal(c1[k], c2[k]), k
    if energy:
        e1, e2 = r1.energy(), r2.energy()
        assert np.array_equal(e1[0], e2[0]) and np.array_equal(e1[1], e2[1])
    assert r1.stats() == r2.stats()


@pytest.mark.parametrize("kind", ["box", "periodic"])
def test_engine_by_engine_equals_whole_steps_and_oracle(hip_ctx, kind):
    ctx = hip_ctx
    a, p, cell = _scene(kind, ctx)
    eng = runner.HipRun(a, p, cell, ctx=ctx)   # one ABI call per engine
    whole = runner.HipRun(a, p, cell, ctx=ctx)  # wdm_step
    o = OL.OracleRun(a, p, cell)
    if kind == "box":  # utils.defaultEngines puts DynDt after ContactLoop (py/utils.py:76-83)
        for r in (eng, whole):
            ctx.set_dyndt(r.sid, make_dyndt(stepPeriod=100, maxRelInc=1e-3))
        o.set_dyndt(make_dyndt(stepPeriod=100, maxRelInc=1e-3))
    m = eng.a.mass
    grav = np.asarray(p.gravity[:])
    free = (eng.a.flags & (soa.GRAVITY_SKIP | soa.CLUMP_CLUMPED)) == 0
    steps = 600 if kind == "box" else 300
    checks = {0, 1, 2, 37, steps // 2, steps - 1}
    done = 0
    for s in range(steps):
        sid = eng.sid
        ctx.step_begin(sid)
        ctx.leapfrog_run(sid)
        if s in checks:
            nd = eng.nodes()  # after Leapfrog: the forces are its reset
            want = np.where(free[:, None], m[:, None] * grav[None, :], 0.)
            assert np.array_equal(nd["force"], want) and not nd["torque"].any()
        ctx.collider_run(sid)
        if s in checks:
            eng.contacts()
        ctx.contact_loop_run(sid)
        ctx.dyndt_engine_run(sid)
        ctx.step_end(sid)
        if s in checks:
            whole.step(s + 1 - done)
            o.step(s + 1 - done)
            done = s + 1
            _same(eng, whole)
            ng, no = eng.nodes(), o.nodes()
            for q in ("pos", "vel", "angVel", "force", "torque"):
                scale = max(np.abs(no[q]).max(), 1e-300)
                assert np.abs(ng[q] - no[q]).max() <= 1e-9 * scale, (s, q)
            assert np.array_equal(OL.pair_keys(eng.contacts()), OL.pair_keys(o.contacts()))
            assert ctx.time(eng.sid) == o.time()
    assert eng.stats()["nFullRuns"] == o.stats()["nFullRuns"] > 1
    eng.close()
    whole.close()


def test_second_leapfrog_reads_the_reset(hip_ctx):
    """Two Leapfrog runs with no ContactLoop between them: the second integrates the reset
    force m g (what DemData.force holds after the first), not the contact sums."""
    ctx = hip_ctx
    a = scenes.box_lattice(nx=5, ny=4, nz=4)
    p, cell = runner.params_from_meta(a), runner.cell_from_meta(a)
    r = runner.HipRun(a, p, cell, ctx=ctx)
    r.step(2600)  # settled: real contacts carry the spheres
    assert r.contacts()["isReal"].sum() > 10
    v0 = r.nodes()["vel"]
    ctx.step_begin(r.sid)
    ctx.leapfrog_run(r.sid)
    v1 = r.nodes()["vel"]
    ctx.leapfrog_run(r.sid)
    v2 = r.nodes()["vel"]
    sph = np.nonzero(a.shape == soa.SHAPE_SPHERE)[0]
    dt, g, d = p.dt, p.gravity[2], p.damping
    assert not np.array_equal(v1[sph] - v0[sph], v2[sph] - v1[sph])  # the first one saw contact forces
    # the second: a = g with the local damping's sign rule (Leapfrog.cpp:32-34)
    dv = v2[sph, 2] - v1[sph, 2]
    sgn = np.sign(g * (v1[sph, 2] + .5 * dt * g))
    assert np.allclose(dv, g * (1 - d * sgn) * dt, rtol=1e-12, atol=0)
    r.close()
