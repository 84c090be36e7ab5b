"""The full-run contact-set checker (tests/bound_pairs.py) pinned against the oracle's
collider on periodic packings: right after a full run the contact list holds exactly the
pairs whose Verlet-enlarged bounds overlap in the periodic cell, each with the periods of
spatialOverlapPeri (pkg/dem/InsertionSortCollider.cpp:903-957) as its cellDist; the
checker misses a halved Verlet distance or a cell one part in 10^4 too large."""
import numpy as np

import bound_pairs as BP
import oracle_lib as OL
from woodem_amd import runner, scenes

def test_checker_equals_oracle_full_runs():
    a = scenes.periodic_hertz(n=2500, seed=4, velStd=.5)
    p = runner.params_from_meta(a)
    cell = runner.cell_from_meta(a)
    o = OL.OracleRun(a, p, cell)
    r = a.radius
    seen = 0
    prev = 0
    for s in range(400):
        o.step(1)
        st = o.stats()
        if st["nFullRuns"] == prev:
            continue
        prev = st["nFullRuns"]
        pos = o.nodes()["pos"]
        size = BP.cell_size(o.cell()[0])
        vd = st["verletDist"]
        c = o.contacts()
        want = BP.periodic_bound_pairs(pos, r, p.distFactor, vd, size)
        assert np.array_equal(OL.pair_keys(c), want), "full run %d at step %d" % (prev, st["step"])
        ok, per = BP.contact_periods(pos, r, p.distFactor, vd, size, c["idA"], c["idB"])
        assert ok.all()
        assert np.array_equal(per, c["cellDist"].astype(np.int64))
        assert (np.abs(per).max(axis=1) > 0).sum() > 50  # contacts across the boundary
        # plausible slips change the set
        assert not np.array_equal(BP.periodic_bound_pairs(pos, r, p.distFactor, vd / 2, size), want)
        assert not np.array_equal(BP.periodic_bound_pairs(pos, r, p.distFactor, vd, size * (1 + 1e-4)), want)
        seen += 1
        if seen == 3:
            break
    assert seen == 3
    o.close()
