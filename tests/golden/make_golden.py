"""Writes the golden fixtures of this directory from the CPU oracle (run from the repository root:
python tests/golden/make_golden.py).

box_n500.npz   500 spheres (10x10x5 lattice) + 6 walls, IdealElPl, gravity, damping .2, settled
               for 3000 steps first; saved: that state (inputs), forces/torques after 1 step,
               state, contact set and energies after 200 steps.
periodic_n600.npz  600 spheres in a periodic cell, Hertz (en=.7), random overlapping placement,
               cell strain rate 1e-2/s; same outputs after 1 and 200 steps.

The fixtures pin the oracle against regressions and are the reference the HIP path is compared
with in tests/test_golden.py.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle_lib as OL  # noqa: E402
from woodem_amd import runner, scenes, soa  # noqa: E402
from woodem_amd._native import LAW_HERTZ_DMT  # noqa: E402

FIELDS = ["pos", "ori", "vel", "angVel", "mass", "inertia", "flags", "force", "torque", "shape", "nodes", "radius", "axis",
          "sense", "mask", "mat", "density", "young", "ktDivKn", "tanPhi"]
PARAMS = [f for f, _ in OL.OrcParams._fields_ if f != "nThreads"]


def save(path, a, p, cell, o):
    out = {"in_" + f: getattr(a, f) for f in FIELDS}
    for f in PARAMS:
        out["p_" + f] = np.array(list(getattr(p, f))) if f == "gravity" else np.array(getattr(p, f))
    out["cell_periodic"] = np.array(cell["periodic"])
    out["cell_hSize"], out["cell_gradV"], out["cell_nextGradV"] = cell["hSize"], cell["gradV"], cell["nextGradV"]
    out["cell_homoDeform"] = np.array(cell["homoDeform"])
    o.step(1)
    n = o.nodes()
    out["s1_force"], out["s1_torque"] = n["force"], n["torque"]
    out["s1_pairs"] = OL.pair_keys(o.contacts())
    o.step(199)
    n = o.nodes()
    for k in ("pos", "vel", "angVel", "force", "torque"):
        out["s200_" + k] = n[k]
    c = o.contacts()
    out["s200_pairs"] = OL.pair_keys(c)
    out["s200_nreal"] = np.array(int(c["isReal"].sum()))
    e, cr = o.energy()
    out["s200_energy"], out["s200_energy_created"] = e, cr.astype(np.int32)
    np.savez_compressed(path, **out)
    print(path, "contacts", len(c["idA"]), "real", int(c["isReal"].sum()))


def main():
    a = scenes.box_lattice(nx=10, ny=10, nz=5)
    p = runner.params_from_meta(a, trackEnergy=True)
    cell = runner.cell_from_meta(a)
    o = OL.OracleRun(a, p, cell, nThreads=1)
    o.step(3000)
    st = o.nodes()
    b = a.copy()
    b.pos, b.vel, b.angVel, b.ori = st["pos"], st["vel"], st["angVel"], st["ori"]
    save(os.path.join(HERE, "box_n500.npz"), b, p, cell, OL.OracleRun(b, p, cell, nThreads=1))

    rng = np.random.default_rng(7)
    n, L = 600, .025
    r = rng.uniform(.8e-3, 1.4e-3, n)
    c = soa.spheres_arrays(rng.uniform(0, L, (n, 3)), r, density=2650, young=7e8, tanPhi=.4, vel=rng.normal(0, .2, (n, 3)))
    cellp = dict(periodic=True, hSize=np.diag([L, L, L]), gradV=np.zeros((3, 3)), nextGradV=1e-2 * np.eye(3), homoDeform=4)
    pp = runner.make_params(dt=5e-7, law=LAW_HERTZ_DMT, poisson=.3, en=.7, trackEnergy=True)
    save(os.path.join(HERE, "periodic_n600.npz"), c, pp, cellp, OL.OracleRun(c, pp, cellp, nThreads=1))


if __name__ == "__main__":
    main()
