"""Batched stepping: scenes stepped together by one wdm_step call (one launch per engine
for all of them, collider full runs merged into one pass) give bit-identical results to
the same scenes stepped alone, whatever else is in the batch and whenever a scene joins.

Each scene alone is a batch of one, which the parity tests pin against the oracle."""
import numpy as np
import pytest

from woodem_amd import runner, scenes
from woodem_amd._native import PREC_MIXED, Context

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "ori", "vel", "angVel", "force", "torque")


def scene_set(ctx):
    out = []
    a = scenes.box_lattice(nx=6, ny=5, nz=5, seed=11)
    out.append((a, runner.params_from_meta(a, trackEnergy=True), runner.cell_from_meta(a)))
    b = scenes.box_lattice(nx=5, ny=4, nz=7, seed=12, radius=4.5e-3)
    pb = runner.params_from_meta(b, precision=PREC_MIXED)
    pb.damping = .1
    out.append((b, pb, runner.cell_from_meta(b)))
    c = scenes.relax_and_order(scenes.periodic_hertz(n=900, seed=13), ctx, iters=150)
    out.append((c, runner.params_from_meta(c, trackEnergy=True), runner.cell_from_meta(c)))
    d = scenes.facet_floor(n=5, seed=14)
    out.append((d, runner.params_from_meta(d, trackEnergy=True), runner.cell_from_meta(d)))
    e = scenes.relax_box(scenes.deposition_box(seed=15, n=700), ctx, iters=100)
    out.append((e, runner.params_from_meta(e, precision=PREC_MIXED), runner.cell_from_meta(e)))
    f = scenes.drum(seed=16, n=400, omega=2., gap=1e-4)
    ns = f.meta["n_spheres"]
    f.vel[:ns, 2] = -.6
    out.append((f, runner.params_from_meta(f, trackEnergy=True), runner.cell_from_meta(f)))
    return out


def state(ctx, sid, a):
    n = ctx.download_nodes(sid, a.n_node)
    c = ctx.contacts(sid)
    e, cr = ctx.energy(sid)
    return n, c, e, cr, ctx.stats(sid)


def assert_same(x, y, label):
    nx, cx, ex, crx, sx = x
    ny, cy, ey, cry, sy = y
    for f in FIELDS:
        assert np.array_equal(nx[f], ny[f]), "%s: %s differs" % (label, f)
    for f in ("idA", "idB", "cellDist", "isReal", "forceLocal", "contPt", "trsf", "uN"):
        assert np.array_equal(cx[f], cy[f], equal_nan=cx[f].dtype.kind == "f"), "%s: contacts %s differ" % (label, f)
    assert np.array_equal(ex, ey) and np.array_equal(crx, cry), label + ": energy differs"
    assert sx == sy, label + ": stats differ"


def test_batch_equals_alone():
    ctx = Context(0)
    try:
        sc = scene_set(ctx)
        steps = (1, 120, 380)
        # alone
        ref = []
        for a, p, cell in sc:
            g = runner.HipRun(a, p, cell, ctx=ctx)
            snaps = []
            for k in steps:
                g.step(k)
                snaps.append(state(ctx, g.sid, g.a))
            ref.append(snaps)
            g.close()
        # all together
        runs = [runner.HipRun(a, p, cell, ctx=ctx) for a, p, cell in sc]
        sids = [g.sid for g in runs]
        for j, k in enumerate(steps):
            ctx.step(sids, k)
            for i, g in enumerate(runs):
                assert_same(state(ctx, g.sid, g.a), ref[i][j], "scene %d after %d steps" % (i, sum(steps[:j + 1])))
        # full runs after the first one happened, several scenes at once at step 0
        assert sum(ctx.stats(s)["nFullRuns"] > 1 for s in sids) >= 2
        for g in runs:
            g.close()
    finally:
        ctx.close()


def test_scene_joining_a_running_batch():
    """Continuous batching: a scene added to a batch after other scenes already ran
    still follows its own trajectory exactly (scene order in the list does not matter)."""
    ctx = Context(0)
    try:
        sc = scene_set(ctx)
        a, p, cell = sc[0]
        g = runner.HipRun(a, p, cell, ctx=ctx)
        g.step(300)
        ref = state(ctx, g.sid, g.a)
        g.close()
        others = [runner.HipRun(x, q, c, ctx=ctx) for x, q, c in sc[1:]]
        ctx.step([o.sid for o in others], 77)
        g = runner.HipRun(a, p, cell, ctx=ctx)
        ctx.step([o.sid for o in others[::-1]] + [g.sid], 150)
        ctx.step([g.sid, others[1].sid], 150)
        assert_same(state(ctx, g.sid, g.a), ref, "joined scene")
        for o in others + [g]:
            o.close()
    finally:
        ctx.close()


def test_scene_listed_twice_rejected(hip_ctx):
    a = scenes.box_lattice(nx=3, ny=3, nz=3)
    g = runner.HipRun(a, runner.params_from_meta(a), runner.cell_from_meta(a), ctx=hip_ctx)
    with pytest.raises(Exception):
        hip_ctx.step([g.sid, g.sid], 1)
    g.close()
