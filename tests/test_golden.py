"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the oracle):
the oracle must reproduce them exactly (regression pin, CPU) and the HIP path must match them
(GPU: contact sets identical; states within 1e-12 relative after 1 step and 1e-9 after 200)."""
import os

import numpy as np
import pytest

import oracle_lib as OL
from woodem_amd import soa

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = ["box_n500.npz", "periodic_n600.npz"]
FIELDS = ["pos", "ori", "vel", "angVel", "mass", "inertia", "flags", "force", "torque", "shape", "nodes", "radius", "axis",
          "sense", "mask", "mat", "density", "young", "ktDivKn", "tanPhi"]


def load(name):
    z = np.load(os.path.join(HERE, name), allow_pickle=False)
    a = soa.DemArrays(**{f: z["in_" + f] for f in FIELDS}).normalized()
    p = OL.OrcParams()
    for f, _ in OL.OrcParams._fields_:
        if f == "nThreads":
            continue
        v = z["p_" + f]
        if f == "gravity":
            for k in range(3):
                p.gravity[k] = float(v[k])
        else:
            setattr(p, f, v.item())
    p.nThreads = 1
    cell = dict(periodic=bool(z["cell_periodic"]), hSize=z["cell_hSize"], gradV=z["cell_gradV"], nextGradV=z["cell_nextGradV"],
                homoDeform=int(z["cell_homoDeform"]))
    return a, p, cell, z


def rel(a, b):
    return float(np.abs(a - b).max()) / max(float(np.abs(b).max()), 1e-300)


def check(run, z, tol1, tol200):
    run.step(1)
    n = run.nodes()
    assert np.array_equal(OL.pair_keys(run.contacts()), z["s1_pairs"])
    assert rel(n["force"], z["s1_force"]) <= tol1
    assert rel(n["torque"], z["s1_torque"]) <= tol1
    run.step(199)
    n = run.nodes()
    c = run.contacts()
    assert np.array_equal(OL.pair_keys(c), z["s200_pairs"])
    assert int(c["isReal"].sum()) == int(z["s200_nreal"])
    for k in ("pos", "vel", "angVel", "force", "torque"):
        assert rel(n[k], z["s200_" + k]) <= tol200, k
    e, cr = run.energy()
    assert np.array_equal(cr.astype(np.int32), z["s200_energy_created"])
    assert np.allclose(e, z["s200_energy"], rtol=max(tol200, 1e-12), atol=1e-14 * np.abs(z["s200_energy"]).max())


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_reproduces_golden(name):
    a, p, cell, z = load(name)
    check(OL.OracleRun(a, p, cell), z, 0., 0.)


@pytest.mark.gpu
@pytest.mark.parametrize("name", FIXTURES)
def test_hip_matches_golden(name, hip_ctx):
    from woodem_amd import runner
    a, p, cell, z = load(name)
    wp = runner.make_params(dt=p.dt)
    for f, _ in OL.OrcParams._fields_:
        if f == "nThreads":
            continue
        if f == "gravity":
            for k in range(3):
                wp.gravity[k] = p.gravity[k]
        else:
            setattr(wp, f, getattr(p, f))
    g = runner.HipRun(a, wp, cell, ctx=hip_ctx)
    check(g, z, 1e-12, 1e-9)
    g.close()
