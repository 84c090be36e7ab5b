"""The collider's two broadphases (multi-level grid, sort and sweep) feed the same exact test,
so a scene stepped with either has the same contact set, in the same order, and the same
trajectory bit for bit: box with walls (IdealElPl), periodic Hertz packing with contacts across
the cell boundary, and a rotating drum of facets (large bounds)."""
import numpy as np
import pytest

import oracle_lib as OL
from woodem_amd import runner, scenes

pytestmark = pytest.mark.gpu


def _run(ctx, a, mode, chunks):
    ctx.set_broadphase(mode)
    r = runner.HipRun(a, runner.params_from_meta(a), runner.cell_from_meta(a), ctx=ctx)
    out = []
    for n in chunks:
        ctx.set_broadphase(mode)
        r.step(n)
        c = r.contacts()
        nd = r.nodes()
        out.append((OL.pair_keys(c), c["cellDist"].copy(), c["isReal"].copy(), nd["pos"].copy(), nd["vel"].copy(),
                    nd["force"].copy(), r.stats()["nFullRuns"]))
    r.close()
    return out


@pytest.mark.parametrize("kind", ["box", "periodic", "drum"])
def test_sweep_equals_grid(hip_ctx, kind):
    if kind == "box":
        a = scenes.box_lattice(nx=10, ny=8, nz=6)
        chunks = (1, 300, 700)
    elif kind == "periodic":
        a = scenes.relax_and_order(scenes.periodic_hertz(n=3000, seed=3), hip_ctx, iters=100)
        chunks = (1, 200, 400)
    else:
        a = scenes.drum(seed=2, n=2000)
        chunks = (1, 300, 500)
    try:
        grid = _run(hip_ctx, a, 0, chunks)
        sweep = _run(hip_ctx, a, 1, chunks)
    finally:
        hip_ctx.set_broadphase(0)
    assert grid[-1][6] >= 2  # full runs after the first
    for k, (g, s) in enumerate(zip(grid, sweep)):
        assert np.array_equal(g[0], s[0]), "chunk %d: contact keys differ" % k
        for q in range(1, 6):
            assert np.array_equal(g[q], s[q]), "chunk %d: field %d differs" % (k, q)
        assert g[6] == s[6]
    if kind == "periodic":
        cd = grid[-1][1]
        assert np.any(cd != 0)  # contacts across the periodic boundary
