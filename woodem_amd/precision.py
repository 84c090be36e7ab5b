"""Mixed-precision error report: FP32 contact geometry and law on FP64 state, against the FP64
path from the same state.

The scene is the 10,000-sphere box (scenes.box_lattice: 25x20x20 lattice at 11 mm, r = 5 mm,
6 walls, gravity, damping .2, IdealElPl) settled for `settle` FP64 steps, so that most spheres
rest on others.  From that state both precisions step once, then on to `long_steps`, and the
report gives, after each, per quantity:

* force on the spheres: max |dF| / max |F| and mean |dF| / mean |F|; torque the same with
  the largest torque a contact force can exert (max |F| times the sphere radius) as scale
  (torques of resting spheres are near zero, so their own maximum is no scale);
* the contact set (potential pairs) and the real contacts: equal or not, and the fraction of
  pairs whose realness agrees;
* the energy balance of each run (EnergyTracker sum over its slots, tracked from the settled
  state): its change between step 1 and the last step, relative to the sum of |slot|, and the
  difference of the two runs' slots relative to the same scale.
"""
import numpy as np

from . import runner, scenes
from ._native import PREC_FP64, PREC_MIXED


def _pairs(c):
    return (c["idA"].astype(np.int64) << 32) | c["idB"].astype(np.int64)


def _rel(a, b, rows, scale=None):
    d = np.linalg.norm(a[rows] - b[rows], axis=1)
    m = np.linalg.norm(b[rows], axis=1)
    smax = max(m.max(), 1e-300) if scale is None else scale
    smean = max(m.mean(), 1e-300) if scale is None else scale
    return float(d.max() / smax), float(d.mean() / smean)


def _compare(g64, g32, spheres, radius):
    n64, n32 = g64.nodes(), g32.nodes()
    c64, c32 = g64.contacts(), g32.contacts()
    fmax, fmean = _rel(n32["force"], n64["force"], spheres)
    tscale = float(np.linalg.norm(n64["force"][spheres], axis=1).max() * radius)
    tmax, tmean = _rel(n32["torque"], n64["torque"], spheres, scale=max(tscale, 1e-300))
    k64, k32 = _pairs(c64), _pairs(c32)
    same = bool(np.array_equal(k64, k32))
    common, i64, i32 = np.intersect1d(k64, k32, return_indices=True)
    agree = float(np.mean(c64["isReal"][i64] == c32["isReal"][i32])) if len(common) else 1.
    return dict(force_rel_max=fmax, force_rel_mean=fmean, torque_rel_max=tmax, torque_rel_mean=tmean, torque_scale=tscale,
                contact_set_equal=same, contacts_fp64=int(len(k64)), contacts_mixed=int(len(k32)),
                real_fp64=int(c64["isReal"].sum()), real_mixed=int(c32["isReal"].sum()), realness_agreement=agree)


def mixed_error_report(ctx, settle=6000, long_steps=1000, seed=1):
    a = scenes.box_lattice(seed=seed)
    cell = runner.cell_from_meta(a)
    g = runner.HipRun(a, runner.params_from_meta(a), cell, ctx=ctx)
    g.step(settle)
    st = g.nodes()
    g.close()
    b = a.copy()
    b.pos, b.vel, b.angVel, b.ori = st["pos"], st["vel"], st["angVel"], st["ori"]
    spheres = np.arange(b.meta["n_spheres"])
    radius = float(b.radius[0])
    g64 = runner.HipRun(b, runner.params_from_meta(b, precision=PREC_FP64, trackEnergy=True), cell, ctx=ctx)
    g32 = runner.HipRun(b, runner.params_from_meta(b, precision=PREC_MIXED, trackEnergy=True), cell, ctx=ctx)
    try:
        out = {"scene": "box_lattice 25x20x20 (10,000 spheres, 6 walls), settled %d FP64 steps" % settle}
        g64.step(1)
        g32.step(1)
        out["step_1"] = _compare(g64, g32, spheres, radius)
        e64_1, e32_1 = g64.energy()[0], g32.energy()[0]
        g64.step(long_steps - 1)
        g32.step(long_steps - 1)
        out["step_%d" % long_steps] = _compare(g64, g32, spheres, radius)
        e64, e32 = g64.energy()[0], g32.energy()[0]
        scale = float(np.abs(e64).sum())
        out["energy"] = dict(balance_drift_fp64=float((e64.sum() - e64_1.sum()) / scale),
                             balance_drift_mixed=float((e32.sum() - e32_1.sum()) / scale),
                             slots_rel_diff=float(np.abs(e32 - e64).sum() / scale))
        return out
    finally:
        g64.close()
        g32.close()
