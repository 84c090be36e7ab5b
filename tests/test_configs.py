"""The benchmark configurations against the oracle, at their own sizes.

* cfg1 (10,000 spheres r 5 mm in a 0.3 x 0.25 x 0.25 m box of 6 walls, IdealElPl, FP64,
  2,000 steps): contact set, forces and torques after one step within 1e-12, the state after
  200 and 2,000 steps within 1e-9, energies within 1e-9.
* cfg2 (periodic Hertz, the bench scene's generator) at 100,000 spheres: the contact set after
  the first collider run bit for bit (pairs, functor order, cellDist, realness) with forces
  and torques within 1e-12, and after 200 steps with the state within 1e-9.
* cfg4 (deposition boxes of the 512-scene sweep, seeds 1000..1007: N = round(exp(U(ln 500,
  ln 2e4))), r ~ U[2, 6] mm, tanPhi ~ U[.1, .8], damping ~ U[.05, .4]): stepped as one FP64
  batch, each scene equal to the oracle (contact set exact, 200-step state within 1e-9); and
  the whole sweep of 512 scenes (2.7 M spheres) in the workload's precision (FP32 contact law)
  through its properties: in-batch equals alone, contact sets at the first collider run equal
  to the FP64 batch's, energy balance.
"""
import numpy as np
import pytest

import oracle_lib as OL
from woodem_amd import runner, scenes
from woodem_amd._native import PREC_MIXED

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "angVel", "force", "torque")


def rel_err(a, b):
    return float(np.abs(a - b).max()) / max(float(np.abs(b).max()), 1e-300)


def same_contacts(cg, co):
    """pairs, functor order, cellDist and realness (both lists sorted by (min id, max id))"""
    for c in (cg, co):
        o = np.lexsort((np.maximum(c["idA"], c["idB"]), np.minimum(c["idA"], c["idB"])))
        for k in ("idA", "idB", "cellDist", "isReal"):
            c[k] = c[k][o]
    return all(np.array_equal(cg[k], co[k]) for k in ("idA", "idB", "cellDist", "isReal"))


def compare(g, o, tol, label, energy=True):
    assert same_contacts(g.contacts(), o.contacts()), label + ": contact set differs"
    ng, no = g.nodes(), o.nodes()
    for q in FIELDS:
        e = rel_err(ng[q], no[q])
        assert e <= tol, "%s: %s rel err %g > %g" % (label, q, e, tol)
    if energy:
        vg, cg = g.energy()
        vo, co = o.energy()
        assert np.array_equal(cg, co)
        # energies are sums in the kernels' block order, the oracle's in contact order
        assert np.allclose(vg, vo, rtol=1e-9, atol=1e-12 * np.abs(vo).max()), label


def test_cfg1_box_full_size(hip_ctx):
    a = scenes.box_lattice()  # 25 x 20 x 20 lattice, spacing 11 mm, jitter 0.4 mm, seed 1
    assert a.n_par == 10006
    p = runner.params_from_meta(a, trackEnergy=True)
    cell = runner.cell_from_meta(a)
    g = runner.HipRun(a, p, cell, ctx=hip_ctx)
    o = OL.OracleRun(a, p, cell)
    for n, tol in ((1, 1e-12), (199, 1e-9), (1800, 1e-9)):
        g.step(n)
        o.step(n)
        compare(g, o, tol, "cfg1 step %d" % g.stats()["step"])
    assert g.stats()["nFullRuns"] == o.stats()["nFullRuns"] > 2
    assert o.contacts()["isReal"].sum() > 300
    g.close()


def test_cfg2_workload_at_100k(hip_ctx):
    """The cfg2 workload (same generator, PSD, packing fraction, device relaxation, Morton
    numbering) at 100,000 spheres: the oracle's initial sort and sweep along one axis costs
    N x (spheres overlapping a bound along that axis), ~N^(5/3), which takes tens of minutes at
    1,000,000 on one core (18 s at 100,000, 150 s at 300,000 from unrelaxed positions)."""
    b = scenes.relax_and_order(scenes.periodic_hertz(n=100_000, seed=3), hip_ctx, iters=500)
    p = runner.params_from_meta(b, trackEnergy=True)
    cell = runner.cell_from_meta(b)
    g = runner.HipRun(b, p, cell, ctx=hip_ctx)
    o = OL.OracleRun(b, p, cell)
    g.step(1)
    o.step(1)
    cg = g.contacts()
    assert len(cg["idA"]) > 400_000 and cg["isReal"].sum() > 100_000
    assert (np.abs(cg["cellDist"]).max(axis=1) > 0).sum() > 1000  # contacts across the periodic boundary
    compare(g, o, 1e-12, "cfg2 at 100k, step 1")
    g.step(199)
    o.step(199)
    compare(g, o, 1e-9, "cfg2 at 100k, step 200")
    assert g.stats()["nFullRuns"] == o.stats()["nFullRuns"] >= 1
    g.close()
    o.close()


def _boxes(ctx, seeds, iters=200):
    return [scenes.relax_box(scenes.deposition_box(seed=s), ctx, iters=iters) for s in seeds]


def test_cfg4_boxes_batch_equals_oracle(hip_ctx):
    arrs = _boxes(hip_ctx, range(1000, 1008))
    runs = []
    for a in arrs:
        p = runner.params_from_meta(a, trackEnergy=True)
        runs.append((runner.HipRun(a, p, runner.cell_from_meta(a), ctx=hip_ctx), p))
    sids = [r.sid for r, _ in runs]
    assert sum(a.n_par for a in arrs) > 10000
    hip_ctx.step(sids, 1)
    hip_ctx.step(sids, 199)
    for a, (g, p) in zip(arrs, runs):
        o = OL.OracleRun(a, p, runner.cell_from_meta(a))
        o.step(1)
        o.step(199)
        compare(g, o, 1e-9, "cfg4 seed %d (%d spheres)" % (a.meta["seed"], a.meta["n_spheres"]))
        assert g.stats()["nFullRuns"] == o.stats()["nFullRuns"]
        o.close()
    for g, _ in runs:
        g.close()


def test_cfg4_full_sweep_mixed_properties(hip_ctx):
    """512 scenes in one mixed-precision batch.  Properties: the same scenes alone give the
    same bits (sample of 4); each scene's contact set after its first collider run equals the
    FP64 batch's (the collider works on FP64 positions, identical before the first forces);
    energy stays balanced."""
    ctx = hip_ctx
    arrs = _boxes(ctx, range(1000, 1512))  # the sweep as the batch benchmark builds it (200 relaxation sweeps)
    ntot = sum(a.n_par for a in arrs)
    assert len(arrs) == 512 and ntot > 2_000_000
    mk = lambda a, prec: runner.HipRun(a, runner.params_from_meta(a, trackEnergy=True, precision=prec),
                                       runner.cell_from_meta(a), ctx=ctx)
    mixed = [mk(a, PREC_MIXED) for a in arrs]
    sids = [r.sid for r in mixed]
    ctx.step(sids, 1)
    # the overlaps the relaxation leaves hold elastic energy at the first step; from then
    # on the sum of the tracked terms stays at that value (what is stored is kinetic, elastic,
    # or dissipated by friction and damping, or work of gravity)
    e1 = sum(float(r.energy()[0].sum()) for r in mixed)
    pick = [0, 101, 333, 511]
    first = {i: mixed[i].contacts() for i in pick}
    for i in pick:  # the FP64 scene alone after one step: same contact set
        r = mk(arrs[i], 0)
        r.step(1)
        c = r.contacts()
        r.close()
        assert same_contacts(dict(first[i]), c), i
    ctx.step(sids, 149)
    for i in pick:  # alone in the same precision: the same bits as in the batch
        r = mk(arrs[i], PREC_MIXED)
        r.step(150)
        na, nb = r.nodes(), mixed[i].nodes()
        for q in ("pos", "ori", "vel", "angVel", "force", "torque"):
            assert np.array_equal(na[q], nb[q]), (i, q)
        assert np.array_equal(r.energy()[0], mixed[i].energy()[0])
        r.close()
    tot = np.zeros(8)
    mag = 0.
    for r in mixed:
        v, _ = r.energy()
        tot += v
        mag += np.abs(v).sum()
    assert mag > 0 and e1 > 0 and abs(tot.sum() - e1) / mag < 2e-2
    nreal = sum(int(r.contacts()["isReal"].sum()) for r in mixed[:64])
    assert nreal > 1000
    for r in mixed:
        r.close()
