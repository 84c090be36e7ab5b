"""The 1,000,000-sphere periodic Hertz workload (the benchmark's scene: same generator, PSD,
packing fraction, device relaxation and Morton numbering) checked at its full size through
properties that need no oracle run (the oracle's initial sort takes tens of minutes at this
size; the workload is compared with the oracle at 100,000 spheres in tests/test_configs.py):

* right after each collider full run, read between wdm_collider_run and wdm_contact_loop_run,
  the contact set equals the set of sphere pairs whose Verlet-enlarged bounds overlap in the
  periodic cell, every contact with spatialOverlapPeri's periods as cellDist
  (pkg/dem/InsertionSortCollider.cpp:903-957; tests/bound_pairs.py, pinned against the oracle's
  collider in tests/test_bound_pairs.py);
* with the FP64 and the mixed-precision contact law (the bounds are FP64 in both), and with
  the work list relaid out between the full runs;
* after each of those steps the contact forces cancel: without gravity the node forces sum to
  zero up to rounding (ContactLoop applies each contact's force to both particles,
  pkg/dem/ContactLoop.cpp:188-190, with sign +1 on A and -1 on B, pkg/dem/Contact.cpp:24-35)."""
import numpy as np
import pytest

import bound_pairs as BP
import oracle_lib as OL
from woodem_amd import runner, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_full_size_contact_sets_and_force_balance(hip_ctx, prec):
    from woodem_amd._native import PREC_FP64, PREC_MIXED
    ctx = hip_ctx
    b = scenes.relax_and_order(scenes.periodic_hertz(n=1_000_000), ctx, iters=500)
    p = runner.params_from_meta(b, precision=PREC_MIXED if prec == "mixed" else PREC_FP64)
    run = runner.HipRun(b, p, runner.cell_from_meta(b), ctx=ctx)
    sid = run.sid
    r = run.a.radius
    checked, prev = [], 0
    for s in range(600):
        ctx.step_begin(sid)
        ctx.leapfrog_run(sid)
        ctx.collider_run(sid)
        st = run.stats()
        full = st["nFullRuns"] != prev
        prev = st["nFullRuns"]
        if full:
            pos = run.nodes()["pos"]
            size = BP.cell_size(run.cell()[0])
            c = run.contacts()
            want = BP.periodic_bound_pairs(pos, r, p.distFactor, st["verletDist"], size)
            got = OL.pair_keys(c)
            assert len(got) > 4_000_000
            assert np.array_equal(got, want), "full run %d at step %d: %d contacts, %d overlapping bounds" % (
                prev, s + 1, len(got), len(want))
            ok, per = BP.contact_periods(pos, r, p.distFactor, st["verletDist"], size, c["idA"], c["idB"])
            assert ok.all() and np.array_equal(per, c["cellDist"].astype(np.int64))
            assert (np.abs(per).max(axis=1) > 0).sum() > 50_000
            checked.append(s + 1)
        ctx.contact_loop_run(sid)
        ctx.step_end(sid)
        if full:
            f = run.nodes()["force"]
            assert np.abs(f.sum(axis=0)).max() <= 1e-9 * np.abs(f).sum(axis=0).max()
            assert run.contacts()["isReal"].sum() > 1_000_000
        if len(checked) == 2:
            break
    assert checked[0] == 1 and len(checked) == 2, checked
    # the work list was relaid out between the two full runs (every 32 steps by default)
    assert checked[1] - checked[0] < 64 or ctx.relayout(sid)[2] >= 1
    run.close()
