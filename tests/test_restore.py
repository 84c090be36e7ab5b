"""Checkpoint restore against the oracle.

The oracle takes a checkpoint's real contacts (orc_set_contacts: ids, cellDist, trsf, local
force, phys, lens, contact point) and its step and time, as woodem_amd.snapshot.load gives
them to the GPU scene (wdm_upload_contacts).

CPU: an oracle restored from the state of another oracle run continues bit for bit like that
run (positions, velocities, orientations, forces, torques, the real contacts with their local
forces and trsf): a box with walls (IdealElPl) and a periodic Hertz packing.  A mistake in the
restore (a phys field in the wrong slot, a transposed trsf, the contact point of the wrong
side, a missed full collider run) breaks the continuation.
GPU: a GPU checkpoint restored into a new GPU scene and into the oracle steps on identically
in both.
"""
import numpy as np
import pytest

import oracle_lib as OL
from woodem_amd import runner, scenes, snapshot


def _state_arrays(a, nodes):
    """the scene arrays with the nodes' state of a run"""
    b = a.copy()
    for k in ("pos", "ori", "vel", "angVel", "force", "torque"):
        setattr(b, k, np.array(nodes[k], dtype=np.float64))
    return b


def _restore_oracle(a, params, cell, con, step, time):
    o = OL.OracleRun(a, params, cell)
    o.set_time(step, time)
    o.set_contacts(con)
    return o


def _real_contacts(o):
    c = o.contacts()
    ph, ln = o.contact_phys()
    r = c["isReal"]
    return dict(idA=c["idA"][r], idB=c["idB"][r], cellDist=c["cellDist"][r], trsf=c["trsf"][r].reshape(-1, 9),
                forceLocal=c["forceLocal"][r], phys=ph[r], lens=ln[r], contPt=c["contPt"][r])


def _compare(o1, o2):
    n1, n2 = o1.nodes(), o2.nodes()
    for q in ("pos", "vel", "angVel", "ori", "force", "torque"):
        assert np.array_equal(n1[q], n2[q]), q
    c1, c2 = o1.contacts(), o2.contacts()
    r1, r2 = c1["isReal"], c2["isReal"]
    assert np.array_equal(OL.pair_keys({k: v[r1] for k, v in c1.items()}), OL.pair_keys({k: v[r2] for k, v in c2.items()}))
    assert np.array_equal(c1["forceLocal"][r1], c2["forceLocal"][r2])
    assert np.array_equal(c1["trsf"][r1], c2["trsf"][r2])


def _cpu_scene(kind):
    if kind == "box":
        a = scenes.box_lattice(nx=6, ny=5, nz=4)
        return a, runner.params_from_meta(a), runner.cell_from_meta(a), 2600
    a = scenes.periodic_hertz(n=300, seed=5)
    return a, runner.params_from_meta(a), runner.cell_from_meta(a), 150


@pytest.mark.parametrize("kind", ["box", "periodic"])
def test_oracle_restore_continues_like_the_uninterrupted_run(kind):
    a, p, cell, n0 = _cpu_scene(kind)
    o1 = OL.OracleRun(a, p, cell)
    o1.step(n0)
    con = _real_contacts(o1)
    assert len(con["idA"]) > 10
    h, g = o1.cell()
    cell2 = dict(cell, hSize=h, gradV=g)  # the cell as a checkpoint holds it (nextGradV unchanged)
    t, _ = o1.time()
    o2 = _restore_oracle(_state_arrays(a, o1.nodes()), p, cell2, con, o1.stats()["step"], t)
    for n in (1, 17, 200):
        o1.step(n)
        o2.step(n)
        _compare(o1, o2)
        assert o1.stats()["step"] == o2.stats()["step"]


@pytest.mark.parametrize("mistake", ["phys", "trsf", "contPt"])
def test_oracle_restore_with_a_planted_mistake_diverges(mistake):
    # the continuation check sees each of these: kn and kt exchanged, trsf transposed, the
    # contact point moved by a micrometre
    a, p, cell, n0 = _cpu_scene("box")
    o1 = OL.OracleRun(a, p, cell)
    o1.step(n0)
    con = _real_contacts(o1)
    if mistake == "phys":
        con["phys"][:, [0, 1]] = con["phys"][:, [1, 0]]
    elif mistake == "trsf":
        con["trsf"] = con["trsf"].reshape(-1, 3, 3).transpose(0, 2, 1).reshape(-1, 9).copy()
    else:
        con["contPt"] = con["contPt"] + 1e-6
    h, g = o1.cell()
    t, _ = o1.time()
    o2 = _restore_oracle(_state_arrays(a, o1.nodes()), p, dict(cell, hSize=h, gradV=g), con, o1.stats()["step"], t)
    o1.step(200)
    o2.step(200)
    with pytest.raises(AssertionError):
        _compare(o1, o2)


def test_oracle_restore_rejects_bad_contacts():
    a, p, cell, _ = _cpu_scene("box")
    o = OL.OracleRun(a, p, cell)
    con = dict(idA=np.array([0], np.int32), idB=np.array([0], np.int32), cellDist=np.zeros((1, 3), np.int32),
               trsf=np.eye(3).reshape(1, 9), forceLocal=np.zeros((1, 3)), phys=np.zeros((1, 5)), lens=np.ones((1, 2)),
               contPt=np.zeros((1, 3)))
    with pytest.raises(RuntimeError):
        o.set_contacts(con)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["box", "periodic"])
def test_gpu_checkpoint_restores_like_the_oracle(kind, tmp_path):
    from woodem_amd._native import Context
    ctx = Context(0)
    if kind == "box":
        a = scenes.box_lattice(nx=8, ny=6, nz=5)
        n0 = 2700
    else:
        a = scenes.relax_and_order(scenes.periodic_hertz(n=3000, seed=4), ctx, iters=100)
        n0 = 300
    p, cell = runner.params_from_meta(a), runner.cell_from_meta(a)
    r1 = runner.HipRun(a, p, cell, ctx=ctx)
    r1.step(n0)
    f = snapshot.save(r1, str(tmp_path / ("%s.npz" % kind)))
    r1.close()
    r2 = snapshot.load(f, ctx=ctx)
    b, meta, cell2, con, _ = snapshot.read(f)
    assert len(con["idA"]) > 10
    o = _restore_oracle(b, p, cell2, con, meta["step"], meta["time"])
    for n in (1, 37, 300):
        r2.step(n)
        o.step(n)
        _compare(r2, o)
    r2.close()
