"""Mixed precision (L6Geom, Cp2 and Law2 in FP32 on FP64 positions, velocities and Leapfrog)
against the FP64 path on the same state.  Bars: contact set identical after one step; one-step
force error max |dF| / max |F| <= 1e-5 (the workload's tolerance); torque and energies reported."""
import numpy as np
import pytest

import oracle_lib as OL
from woodem_amd import runner, scenes
from woodem_amd._native import PREC_FP64, PREC_MIXED

pytestmark = pytest.mark.gpu


def settled_box(hip_ctx, steps=2500):
    a = scenes.box_lattice(nx=10, ny=8, nz=6)
    p = runner.params_from_meta(a)
    g = runner.HipRun(a, p, runner.cell_from_meta(a), ctx=hip_ctx)
    g.step(steps)
    st = g.nodes()
    g.close()
    b = a.copy()
    b.pos, b.vel, b.angVel, b.ori = st["pos"], st["vel"], st["angVel"], st["ori"]
    return b


@pytest.mark.parametrize("law", ["IdealElPl", "Hertz"])
def test_mixed_one_step_force_error(hip_ctx, law):
    if law == "Hertz":
        a = scenes.periodic_hertz(n=4000)
        b = scenes.relax_and_order(a, hip_ctx, iters=200)
    else:
        b = settled_box(hip_ctx)
    cell = runner.cell_from_meta(b)
    g64 = runner.HipRun(b, runner.params_from_meta(b, precision=PREC_FP64), cell, ctx=hip_ctx)
    g32 = runner.HipRun(b, runner.params_from_meta(b, precision=PREC_MIXED), cell, ctx=hip_ctx)
    g64.step(1)
    g32.step(1)
    c64, c32 = g64.contacts(), g32.contacts()
    assert np.array_equal(OL.pair_keys(c64), OL.pair_keys(c32))
    assert np.array_equal(c64["isReal"], c32["isReal"]) and c64["isReal"].sum() > 50
    n64, n32 = g64.nodes(), g32.nodes()
    F64, F32 = n64["force"], n32["force"]
    err = np.abs(F32 - F64).max() / np.abs(F64).max()
    assert err <= 1e-5, err
    g64.close()
    g32.close()


def test_mixed_long_run_statistics(hip_ctx):
    """After 1000 steps the mixed run keeps the same number of contacts and energy within 1%."""
    a = scenes.box_lattice(nx=10, ny=8, nz=6)
    cell = runner.cell_from_meta(a)
    g64 = runner.HipRun(a, runner.params_from_meta(a, trackEnergy=True, precision=PREC_FP64), cell, ctx=hip_ctx)
    g32 = runner.HipRun(a, runner.params_from_meta(a, trackEnergy=True, precision=PREC_MIXED), cell, ctx=hip_ctx)
    g64.step(3000)
    g32.step(3000)
    e64, _ = g64.energy()
    e32, _ = g32.energy()
    s = np.abs(e64).sum()
    assert np.abs(e64 - e32).sum() / s < 1e-2
    n64 = g64.contacts()["isReal"].sum()
    n32 = g32.contacts()["isReal"].sum()
    assert abs(int(n64) - int(n32)) <= max(3, .05 * n64)
    g64.close()
    g32.close()


def test_mixed_error_report(hip_ctx):
    """The report bench.py carries for the mixed-precision line (config-1 scene, 1 and 1000
    steps): one-step force error within the workload's 1e-5, same contact set at one step."""
    from woodem_amd.precision import mixed_error_report
    r = mixed_error_report(hip_ctx, settle=6000, long_steps=200)
    s1 = r["step_1"]
    assert s1["contact_set_equal"] and s1["real_fp64"] > 1000
    assert s1["force_rel_max"] <= 1e-5 and s1["force_rel_mean"] <= s1["force_rel_max"] and s1["torque_rel_max"] <= 1e-5
    assert 0 <= r["step_200"]["realness_agreement"] <= 1
    assert abs(r["energy"]["balance_drift_fp64"]) < 1e-2
