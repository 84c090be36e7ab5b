"""Windows of steps: wdm_step launches several steps of a batch back to back (one captured
HIP graph, replayed while the batch's tables keep their shape) and looks at the Verlet
triggers once per window; a scene whose trigger fires stops at that step and finishes it
(collider, ContactLoop) after the window.  The results are bit-identical to stepping one
step at a time, for every window length, with and without graphs, and for any collider
cell width factor; graphs are replayed with downloads (eager work and synchronisations)
between the replays and every replay is checked."""
import numpy as np
import pytest

from woodem_amd import runner, scenes
from woodem_amd._native import Context

from test_gpu_batch import assert_same, scene_set, state

pytestmark = pytest.mark.gpu


def run_set(ctx, sc, chunks):
    runs = [runner.HipRun(a, p, cell, ctx=ctx) for a, p, cell in sc]
    sids = [g.sid for g in runs]
    out = []
    for k in chunks:
        ctx.step(sids, k)
        out.append([state(ctx, g.sid, g.a) for g in runs])
    for g in runs:
        g.close()
    return out


def test_windows_equal_step_by_step():
    ctx = Context(0)
    try:
        sc = scene_set(ctx)
        chunks = (1, 37, 200, 163)
        ctx.set_tuning(window=1, graphs=0, cell_factor=1.)
        ref = run_set(ctx, sc, chunks)
        for window, graphs, cf in ((16, 1, 1.), (5, 0, 1.), (64, 1, 1.7)):
            ctx.set_tuning(window=window, graphs=graphs, cell_factor=cf)
            t0 = ctx.tuning()
            got = run_set(ctx, sc, chunks)
            for j, k in enumerate(chunks):
                for i in range(len(sc)):
                    assert_same(got[j][i], ref[j][i], "window %d graphs %d cells x%g: scene %d after chunk %d" % (window, graphs, cf, i, j))
            t1 = ctx.tuning()
            assert t1["windows"] > t0["windows"]
            if graphs:
                assert t1["graph_launches"] > t0["graph_launches"]
            else:
                assert t1["graph_launches"] == t0["graph_launches"]
    finally:
        ctx.close()


def test_graph_replays_between_full_runs():
    """One strained periodic scene: between two collider runs the window's graph keeps its
    shape and is replayed; each replay is checked against step-by-step stepping."""
    ctx = Context(0)
    try:
        a = scenes.relax_and_order(scenes.periodic_hertz(n=2000, seed=21), ctx, iters=150)
        p = runner.params_from_meta(a, trackEnergy=True)
        cell = runner.cell_from_meta(a)
        chunks = [24] * 12
        ctx.set_tuning(window=1, graphs=0)
        ref = run_set(ctx, [(a, p, cell)], chunks)
        ctx.set_tuning(window=8, graphs=1)
        t0 = ctx.tuning()
        got = run_set(ctx, [(a, p, cell)], chunks)
        t1 = ctx.tuning()
        for j in range(len(chunks)):
            assert_same(got[j][0], ref[j][0], "after chunk %d" % j)
        launches = t1["graph_launches"] - t0["graph_launches"]
        captures = t1["graph_captures"] - t0["graph_captures"]
        assert launches >= 30 and captures < launches, (launches, captures)
        assert ref[-1][0][4]["nFullRuns"] >= 2
    finally:
        ctx.close()


def test_tuning_rejects_bad_values(hip_ctx):
    with pytest.raises(Exception):
        hip_ctx.set_tuning(window=0)
    with pytest.raises(Exception):
        hip_ctx.set_tuning(cell_factor=.5)
    t = hip_ctx.tuning()
    assert t["window"] >= 1 and t["cell_factor"] >= 1.
