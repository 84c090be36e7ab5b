"""Particles removed and inserted in place (DemField::removeParticle, pkg/dem/Particle.cpp:445-493;
ParticleContainer::insert, ParticleContainer.cpp:23-72) and BoxOutlet (pkg/dem/Outlet.cpp:12-110).

The oracle is pinned by the reference's semantics in closed form: a removed particle's
contacts are gone while its partner keeps the force of the step it was removed in; ids of
inserted spheres come from the freed ones in the order they were freed, then past the end;
an inserted sphere has no force in its first step (gravity enters through the force reset),
so its velocity after two steps is g dt; the outlet selects the particles outside its closed
box that match its mask.  The HIP path is compared with the oracle bit for bit through removals,
insertions and a BoxOutlet run from the Woo-style Scene.
"""
import math

import numpy as np
import pytest

import oracle_lib as OL
from woodem_amd import runner, scenes, soa

R, RHO, E = 5e-3, 3000., 1e7


def pair_scene(spare=0, gravity=(0., 0., 0.)):
    pos = np.array([[0, 0, 0], [2 * R - 2e-5, 0, 0]], float)
    a = soa.spheres_arrays(pos, R, density=RHO, young=E)
    if spare:
        a = soa.with_spare(a, spare)
    p = runner.make_params(dt=1e-6, gravity=gravity)
    return a, p, runner.cell_from_meta(a)


def test_removal_drops_contacts_partner_keeps_force():
    a, p, cell = pair_scene()
    o = OL.OracleRun(a, p, cell)
    o.step(1)
    assert o.contacts()["isReal"].sum() == 1
    f0 = o.nodes()["force"][0].copy()
    assert f0[0] < 0  # pushed away from sphere 1
    v0 = o.nodes()["vel"][0].copy()
    o.remove_particles([1])
    assert len(o.contacts()["idA"]) == 0
    assert np.array_equal(o.nodes()["force"][0], f0)
    o.step(1)
    # Leapfrog with no damping: v += dt F/m, the force of the removal step (ContactLoop applied it first)
    m = soa.sphere_mass_inertia(R, RHO)[0]
    assert o.nodes()["vel"][0][0] == v0[0] + p.dt * (f0[0] / m)
    o.step(10)
    assert len(o.contacts()["idA"]) == 0  # the hole never collides
    with pytest.raises(RuntimeError):
        o.remove_particles([1])  # gone already


def test_insert_ids_follow_particle_container():
    a = scenes.box_lattice(nx=4, ny=3, nz=2)
    a = soa.with_spare(a, 3)
    n0 = a.n_par - 3
    p, cell = runner.params_from_meta(a), runner.cell_from_meta(a)
    o = OL.OracleRun(a, p, cell)
    assert list(o.free_ids()[0]) == [] and o.free_ids()[1] == n0
    o.remove_particles([5, 2, 9])
    assert list(o.free_ids()[0]) == [5, 2, 9]
    m, I = soa.sphere_mass_inertia(R, RHO)
    ids = o.insert_spheres(np.tile([.1, .1, .2], (5, 1)) + np.arange(5)[:, None] * [.02, 0, 0], R, m, I, 0, soa.DEFAULT_MOVABLE_MASK)
    assert list(ids) == [5, 2, 9, n0, n0 + 1]  # freed ids first (FIFO), then past the end
    with pytest.raises(RuntimeError):
        o.insert_spheres(np.zeros((2, 3)), R, m, I, 0, soa.DEFAULT_MOVABLE_MASK)  # one slot left


def test_inserted_sphere_starts_without_gravity():
    g = -9.81
    a, p, cell = pair_scene(spare=1, gravity=(0., 0., g))
    o = OL.OracleRun(a, p, cell)
    o.step(3)
    m, I = soa.sphere_mass_inertia(R, RHO)
    (i,) = o.insert_spheres(np.array([[0, .5, 0]]), R, m, I, 0, soa.DEFAULT_MOVABLE_MASK)
    o.step(1)
    assert np.array_equal(o.nodes()["vel"][i], [0., 0., 0.])  # zero force in its first Leapfrog
    o.step(1)
    assert o.nodes()["vel"][i][2] == p.dt * ((m * g) / m)  # then the reset force m g


def test_select_box_semantics_in_python():
    """the outlet's selection rule, restated on arrays: closed box, inside flag, masks"""
    from woodem_amd import dem
    e = dem.BoxOutlet(box=((0, 0, 0), (1, 1, 1)))
    assert e.mask == 0b0100 and e.markMask == 0 and not e.inside  # Outlet.hpp:23-25, DemField.defaultOutletMask
    e = dem.BoxOutlet(stepPeriod=5)
    assert [s for s in range(12) if e._activate(s, None)] == [0, 5, 10]  # stepModulo: step % period == 0
    assert (e.nDone, e.stepLast, e.stepPrev) == (3, 10, 5)
    e = dem.BoxOutlet(stepPeriod=5)
    assert [s for s in range(3, 23) if e._activate(s, None)] == [5, 10, 15, 20]


# ---------------- HIP path against the oracle ----------------
@pytest.mark.gpu
def test_gpu_remove_insert_equals_oracle():
    from woodem_amd._native import Context
    ctx = Context(0)
    a = soa.with_spare(scenes.box_lattice(nx=8, ny=6, nz=5), 40)
    p = runner.params_from_meta(a, trackEnergy=True)
    cell = runner.cell_from_meta(a)
    g = runner.HipRun(a, p, cell, ctx=ctx)
    o = OL.OracleRun(a, p, cell)
    g.step(1500)
    o.step(1500)
    z = o.nodes()["pos"][:, 2]
    sph = np.nonzero(a.shape == soa.SHAPE_SPHERE)[0]
    rm = sph[np.argsort(-z[sph], kind="stable")[:25]]  # the 25 highest spheres, in that order
    ctx.remove_particles(g.sid, rm)
    o.remove_particles(rm)
    g.step(200)
    o.step(200)
    m, I = soa.sphere_mass_inertia(R, RHO)
    newpos = np.array([[.05 + .012 * k, .12, .2] for k in range(30)])
    ig = ctx.insert_spheres(g.sid, newpos, R, m, I, 0, soa.DEFAULT_MOVABLE_MASK)
    io = o.insert_spheres(newpos, R, m, I, 0, soa.DEFAULT_MOVABLE_MASK)
    assert np.array_equal(ig, io)
    assert np.array_equal(ctx.free_ids(g.sid)[0], o.free_ids()[0]) and ctx.free_ids(g.sid)[1] == o.free_ids()[1]
    for n in (1, 50, 600):
        g.step(n)
        o.step(n)
        ng, no = g.nodes(), o.nodes()
        for q in ("pos", "vel", "angVel", "force", "torque"):
            assert np.array_equal(ng[q], no[q]), q
        assert np.array_equal(OL.pair_keys(g.contacts()), OL.pair_keys(o.contacts()))
    vg, cg = g.energy()
    vo, co = o.energy()
    assert np.array_equal(cg, co) and np.allclose(vg, vo, rtol=1e-9, atol=1e-12 * np.abs(vo).max())
    g.close()


@pytest.mark.gpu
def test_gpu_box_outlet_scene_equals_oracle():
    """spheres falling through an open floor; a BoxOutlet deletes those below it every 20 steps"""
    import woodem_amd as woo
    from woodem_amd import utils
    S = woo.Scene(fields=[woo.DemField(gravity=(0, 0, -9.81))], dt=2e-5, trackEnergy=True)
    walls = [utils.wall(0., axis=0, sense=1), utils.wall(.1, axis=0, sense=-1), utils.wall(0., axis=1, sense=1),
             utils.wall(.1, axis=1, sense=-1)]
    S.dem.par.add(walls)
    S.dem.par.add([utils.sphere((.02 + .015 * (k % 5), .05, .005 + .012 * (k // 5)), .005) for k in range(40)])
    out = woo.BoxOutlet(box=((-1, -1, 0.), (1, 1, 1)), stepPeriod=20, save=True)
    S.engines = utils.defaultEngines(damping=.2, dynDtPeriod=0) + [out]
    a = S.dem._compile()
    p = S._params()
    S.run(1300)
    S.run(1700)  # two calls: the outlet's state carries over
    assert out.num >= 5  # the lowest layer left through the floor
    # the oracle stepped the same way, with the outlet's rule applied to its own state
    o = OL.OracleRun(a, p, runner.cell_from_meta(a))
    mirror = woo.BoxOutlet(box=out.box, stepPeriod=20)
    removed, kin = [], 0.
    for s in range(3000):
        fire = mirror._activate(s, None)
        o.step(1)
        if fire:
            nd = o.nodes()
            pos = nd["pos"][:a.n_par]
            lo, hi = np.array(out.box[0]), np.array(out.box[1])
            inb = np.all((pos >= lo) & (pos <= hi), axis=1)
            live = (a.shape == soa.SHAPE_SPHERE) & ~np.isin(np.arange(a.n_par), removed)
            ids = np.nonzero(live & ~inb & ((a.mask & out.mask) != 0))[0]
            if len(ids):
                v, w = nd["vel"][ids], nd["angVel"][ids]
                kin += float(np.sum(.5 * a.mass[ids] * np.einsum("ij,ij->i", v, v)
                                    + .5 * np.einsum("ij,ij->i", w, a.inertia[ids] * w)))
                o.remove_particles(ids)
                removed.extend(ids.tolist())
    assert sorted(removed) == sorted(i for i in range(a.n_par) if not S.dem.par.exists(i))
    assert out.num == len(removed) and len(out.diamMassTime) == len(removed)
    no = o.nodes()
    for i in range(a.n_par):
        if S.dem.par.exists(i):
            assert np.array_equal(S.dem.par[i].pos, no["pos"][i])
    assert S.energy["kinOutlet"] == pytest.approx(kin, rel=1e-12)


@pytest.mark.gpu
def test_gpu_insert_equals_oracle():
    """spheres inserted into spare slots of a settled box (no removal)"""
    from woodem_amd._native import Context
    ctx = Context(0)
    a = soa.with_spare(scenes.box_lattice(nx=8, ny=6, nz=5), 40)
    p = runner.params_from_meta(a, trackEnergy=True)
    cell = runner.cell_from_meta(a)
    g = runner.HipRun(a, p, cell, ctx=ctx)
    o = OL.OracleRun(a, p, cell)
    g.step(1500)
    o.step(1500)
    m, I = soa.sphere_mass_inertia(R, RHO)
    newpos = np.array([[.05 + .012 * k, .12, .2] for k in range(30)])
    ig = ctx.insert_spheres(g.sid, newpos, R, m, I, 0, soa.DEFAULT_MOVABLE_MASK)
    io = o.insert_spheres(newpos, R, m, I, 0, soa.DEFAULT_MOVABLE_MASK)
    assert np.array_equal(ig, io) and ig[0] == a.n_par - 40
    assert ctx.free_ids(g.sid)[1] == o.free_ids()[1]
    for n in (1, 50, 600):
        g.step(n)
        o.step(n)
        ng, no = g.nodes(), o.nodes()
        for q in ("pos", "vel", "angVel", "force", "torque"):
            assert np.array_equal(ng[q], no[q]), q
        assert np.array_equal(OL.pair_keys(g.contacts()), OL.pair_keys(o.contacts()))
    vg, cg = g.energy()
    vo, co = o.energy()
    assert np.array_equal(cg, co) and np.allclose(vg, vo, rtol=1e-9, atol=1e-12 * np.abs(vo).max())
    g.close()


@pytest.mark.gpu
def test_gpu_box_outlet_marking_equals_oracle():
    """BoxOutlet in marking mode (markMask set): spheres below the box get the mark, nothing is
    removed; marks and trajectory match the oracle with the outlet's rule applied to its state"""
    import woodem_amd as woo
    from woodem_amd import utils
    S = woo.Scene(fields=[woo.DemField(gravity=(0, 0, -9.81))], dt=2e-5)
    walls = [utils.wall(0., axis=0, sense=1), utils.wall(.1, axis=0, sense=-1), utils.wall(0., axis=1, sense=1),
             utils.wall(.1, axis=1, sense=-1)]
    S.dem.par.add(walls)
    S.dem.par.add([utils.sphere((.02 + .015 * (k % 5), .05, .005 + .012 * (k // 5)), .005) for k in range(40)])
    out = woo.BoxOutlet(box=((-1, -1, 0.), (1, 1, 1)), stepPeriod=20, markMask=0b1000, save=True)
    S.engines = utils.defaultEngines(damping=.2, dynDtPeriod=0) + [out]
    a = S.dem._compile()
    mask0 = a.mask.copy()
    p = S._params()
    S.run(1500)
    S.run(1500)
    assert out.num >= 5
    o = OL.OracleRun(a, p, runner.cell_from_meta(a))
    mirror = woo.BoxOutlet(box=out.box, stepPeriod=20)
    mask = mask0.copy()
    marked = []
    for s in range(3000):
        fire = mirror._activate(s, None)
        o.step(1)
        if fire:
            pos = o.nodes()["pos"][:a.n_par]
            inb = np.all((pos >= np.array(out.box[0])) & (pos <= np.array(out.box[1])), axis=1)
            uni = (a.shape == soa.SHAPE_SPHERE) | (a.shape == soa.SHAPE_WALL)
            ids = np.nonzero(uni & ~inb & ((mask & out.mask) != 0) & ((mask & out.markMask) != out.markMask))[0]
            mask[ids] |= out.markMask
            o.set_masks(ids, mask[ids])
            marked.extend(ids.tolist())
    sa = S.dem._arrays  # the scene's arrays, which the outlet updates
    assert out.num == len(marked) and sorted(marked) == sorted(np.nonzero(sa.mask & 0b1000)[0].tolist())
    assert np.array_equal(sa.mask, mask)
    no = o.nodes()
    for i in range(a.n_par):
        assert np.array_equal(S.dem.par[i].pos, no["pos"][i])
