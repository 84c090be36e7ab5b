"""Checkpoints (npz) and VTK output (woodem_amd.snapshot).

CPU: the VTK writer on the oracle's state (through a small adapter with the Context calls it
makes): every array read back from the XML equals the state it was written from, names as
VtkExport's (pkg/dem/VtkExport.cpp:132-260), periodic points wrapped into the cell.
GPU: a checkpoint restored into a new scene steps on bit for bit like the saved scene
(positions, velocities, orientations, forces, real contacts with their local forces and
trsf, step, time, dt, DynDt counters; energies up to the summation order of their block
sums), for a box with walls and energy tracking, the periodic Hertz packing, a rotating
drum with facets and clumps, and mixed precision.
"""
import numpy as np
import pytest

import oracle_lib as OL
from woodem_amd import runner, scenes, snapshot, soa


class OracleCtx:
    """the Context calls vtk_export makes, answered by an OracleRun"""

    def __init__(self, o):
        self.o = o

    def stats(self, sid):
        return self.o.stats()

    def cell(self, sid):
        return self.o.cell()

    def download_nodes(self, sid, n):
        return self.o.nodes()

    def contacts(self, sid):
        return self.o.contacts()


class OracleRunView:
    def __init__(self, a, cell, o):
        self.ctx, self.sid, self.a, self.cell0 = OracleCtx(o), 0, a.normalized(), cell


def _stress_loop(a, c):
    """particleStress (pkg/dem/Funcs.cpp:227-240) contact by contact, other id ascending"""
    sigN, sigT = np.zeros((a.n_par, 3)), np.zeros((a.n_par, 3))
    r = np.nonzero(c["isReal"])[0]
    per = {}
    for k in r:
        per.setdefault(int(c["idA"][k]), []).append((int(c["idB"][k]), k))
        per.setdefault(int(c["idB"][k]), []).append((int(c["idA"][k]), k))
    for p, lst in per.items():
        for _, k in sorted(lst):
            A, B = c["idA"][k], c["idB"][k]
            rA, rB = a.radius[A], a.radius[B]
            rr = min(rA, rB) if (a.shape[A] == soa.SHAPE_SPHERE and a.shape[B] == soa.SHAPE_SPHERE) else rB
            contA = np.pi * (rr * rr)
            f = c["forceLocal"][k]
            sigN[p] += (1 / contA) * (c["trsf"][k] @ np.array([f[0], 0., 0.]))
            sigT[p] += (1 / contA) * (c["trsf"][k] @ np.array([0., f[1], f[2]]))
    return sigN, sigT


def test_vtk_export_box(tmp_path):
    a = scenes.box_lattice(nx=5, ny=4, nz=4)
    p, cell = runner.params_from_meta(a), runner.cell_from_meta(a)
    o = OL.OracleRun(a, p, cell)
    o.step(2600)
    run = OracleRunView(a, cell, o)
    files = snapshot.vtk_export(run, str(tmp_path / "out/box."))
    assert set(files) == {"spheres", "con"}  # no facets: no mesh file
    assert files["spheres"].endswith("box.spheres.2600.vtu") and files["con"].endswith("box.con.2600.vtp")
    s = snapshot.read_vtk_arrays(files["spheres"])
    nd = o.nodes()
    sph = np.nonzero(a.shape == soa.SHAPE_SPHERE)[0]
    assert np.array_equal(s["Points"], nd["pos"][sph])
    assert np.array_equal(s["radius"], a.radius[sph]) and np.array_equal(s["mass"], a.mass[sph])
    assert np.array_equal(s["id"], sph) and np.array_equal(s["vel"], nd["vel"][sph])
    assert np.array_equal(s["angVel"], nd["angVel"][sph]) and np.array_equal(s["matId"], a.mat[sph])
    assert np.array_equal(s["mask"], a.mask[sph]) and np.array_equal(s["color"], np.zeros(len(sph)))
    assert np.array_equal(s["types"], np.ones(len(sph)))  # VTK_VERTEX
    c = o.contacts()
    sN, sT = _stress_loop(a, c)
    assert np.abs(sN).max() > 0 and np.abs(sT).max() > 0
    assert np.allclose(s["sigN"], sN[sph], rtol=1e-12, atol=1e-12 * np.abs(sN).max())
    assert np.allclose(s["sigT"], sT[sph], rtol=1e-12, atol=1e-12 * np.abs(sT).max())
    r = c["isReal"]
    assert r.sum() > 10
    # contacts: every particle's node position at its id, then the contact points added for
    # the wall sides; a line per real contact between the ids (VtkExport.cpp:146-212)
    k = snapshot.read_vtk_arrays(files["con"])
    assert np.array_equal(k["Fn"], c["forceLocal"][r, 0])
    assert np.allclose(k["|Ft|"], np.hypot(c["forceLocal"][r, 1], c["forceLocal"][r, 2]), rtol=1e-15, atol=0)
    pts = k["Points"]
    assert np.array_equal(pts[:a.n_par], nd["pos"][a.nodes[:, 0]])
    ln = k["connectivity"].reshape(-1, 2)
    A, B = c["idA"][r], c["idB"][r]
    wall = a.shape[A] == soa.SHAPE_WALL  # walls come first in a contact
    assert wall.sum() > 0 and (~wall).sum() > 0
    assert np.array_equal(ln[:, 1], B) and np.array_equal(ln[~wall, 0], A[~wall])
    assert np.array_equal(ln[wall, 0], a.n_par + np.arange(wall.sum()))  # added points, in contact order
    assert np.array_equal(pts[ln[wall, 0]], c["contPt"][r][wall])
    assert len(pts) == a.n_par + wall.sum()


def test_vtk_export_periodic_wraps_points(tmp_path):
    a = scenes.periodic_hertz(n=400, seed=3)
    L = a.meta["L"]
    a.pos[:7] += np.array([L, -L, 2 * L])  # outside the cell: exported wrapped into [0, L)
    p, cell = runner.params_from_meta(a), runner.cell_from_meta(a)
    o = OL.OracleRun(a, p, cell)
    run = OracleRunView(a, cell, o)
    files = snapshot.vtk_export(run, str(tmp_path / "p."), what=("spheres",))
    s = snapshot.read_vtk_arrays(files["spheres"])
    assert np.all(s["Points"] >= 0) and np.all(s["Points"] < L)
    h = o.cell()[0]
    want = o.nodes()["pos"] - np.floor(o.nodes()["pos"] / np.diag(h)) * np.diag(h)
    assert np.array_equal(s["Points"], want)


def test_vtk_contacts_across_the_periodic_boundary(tmp_path):
    # a contact whose cellDist differs from the difference of its particles' wrap distances is
    # drawn twice, each time from one particle to the other's shifted copy (VtkExport.cpp:186-201)
    a = scenes.periodic_hertz(n=400, seed=3)
    p, cell = runner.params_from_meta(a), runner.cell_from_meta(a)
    o = OL.OracleRun(a, p, cell)
    o.step(5)
    run = OracleRunView(a, cell, o)
    files = snapshot.vtk_export(run, str(tmp_path / "p."), what=("con",))
    k = snapshot.read_vtk_arrays(files["con"])
    h = o.cell()[0]
    L = np.diag(h)
    pos = o.nodes()["pos"][a.nodes[:, 0]]
    wcd = np.floor(pos / L).astype(np.int64)
    c = o.contacts()
    pts, ln = k["Points"], k["connectivity"].reshape(-1, 2)
    assert np.array_equal(pts[:a.n_par], pos - wcd * L)
    li, extra, crossing = 0, a.n_par, 0
    for j in np.nonzero(c["isReal"])[0]:
        A, B, cd = c["idA"][j], c["idB"][j], c["cellDist"][j]
        if np.array_equal(cd, wcd[B] - wcd[A]):
            assert tuple(ln[li]) == (A, B)
            li += 1
            continue
        crossing += 1
        for me, other in ((A, B), (B, A)):
            assert tuple(ln[li]) == (me, extra)
            assert np.array_equal(pts[extra], pos[me] + h @ (wcd[other] - cd).astype(np.float64))
            li += 1
            extra += 1
    assert li == len(ln) and extra == len(pts) and crossing > 0
    assert len(k["Fn"]) == len(ln)


@pytest.mark.parametrize("halfThick", [0., 1e-3])
def test_vtk_mesh_facets(tmp_path, halfThick):
    # each facet a triangle with its own three points, or with halfThick its two faces and
    # closed flat edges (8 triangles on 6 points, thickFacetDiv=1, VtkExport.cpp:296-313)
    a = scenes.facet_floor(halfThick=halfThick)
    p, cell = runner.params_from_meta(a), runner.cell_from_meta(a)
    o = OL.OracleRun(a, p, cell)
    o.step(3)
    run = OracleRunView(a, cell, o)
    files = snapshot.vtk_export(run, str(tmp_path / "f."), what=("mesh",))
    m = snapshot.read_vtk_arrays(files["mesh"])
    fac = np.nonzero(a.shape == soa.SHAPE_FACET)[0]
    per = 1 if halfThick == 0 else 8
    npt = 3 if halfThick == 0 else 6
    assert len(m["matId"]) == per * len(fac) and len(m["Points"]) == npt * len(fac)
    pos = o.nodes()["pos"]
    tri = m["connectivity"].reshape(-1, 3)
    for i, f in enumerate(fac):
        P = m["Points"][npt * i: npt * (i + 1)]
        A, B, C = pos[a.nodes[f, 0]], pos[a.nodes[f, 1]], pos[a.nodes[f, 2]]
        if halfThick == 0:
            assert np.array_equal(P, np.array([A, B, C]))
        else:
            nn = np.cross(B - A, C - A)
            dz = nn / np.linalg.norm(nn) * halfThick
            assert np.allclose(P, np.array([A + dz, B + dz, C + dz, A - dz, B - dz, C - dz]), rtol=0, atol=1e-15)
        assert np.all((tri[per * i: per * (i + 1)] >= npt * i) & (tri[per * i: per * (i + 1)] < npt * (i + 1)))
    assert np.array_equal(m["matId"], np.repeat(a.mat[fac], per))
    assert np.array_equal(m["vel"], np.repeat(o.nodes()["vel"][a.nodes[fac, 0]], per, axis=0))
    assert np.array_equal(m["|sigma|"], np.zeros(per * len(fac))) and np.array_equal(m["types"], np.full(per * len(fac), 5))


def test_vtk_binary_and_ascii_agree(tmp_path):
    # the inline binary form (base64 of a UInt32 byte count and the little-endian values) holds
    # the same numbers as the ascii form, bit for bit
    a = scenes.box_lattice(nx=5, ny=4, nz=4)
    p, cell = runner.params_from_meta(a), runner.cell_from_meta(a)
    o = OL.OracleRun(a, p, cell)
    o.step(2600)
    run = OracleRunView(a, cell, o)
    fb = snapshot.vtk_export(run, str(tmp_path / "b."))
    fa = snapshot.vtk_export(run, str(tmp_path / "a."), binary=False)
    for part in ("spheres", "con"):
        assert 'format="binary"' in open(fb[part]).read() and 'format="ascii"' in open(fa[part]).read()
        b, c = snapshot.read_vtk_arrays(fb[part]), snapshot.read_vtk_arrays(fa[part])
        assert sorted(b) == sorted(c)
        for k in c:
            assert np.array_equal(b[k], c[k]), (part, k)


def test_checkpoint_rejects_other_files(tmp_path):
    f = tmp_path / "x.npz"
    np.savez(f, meta=np.array('{"schema": "something else"}'))
    with pytest.raises(ValueError):
        snapshot.read(str(f))


# ---------------- GPU: continuation after a restore ----------------
def _compare(r1, r2):
    n1, n2 = r1.nodes(), r2.nodes()
    for q in ("pos", "vel", "angVel", "ori", "force", "torque"):
        assert np.array_equal(n1[q], n2[q]), q
    c1, c2 = r1.contacts(), r2.contacts()
    k1 = OL.pair_keys({k: v[c1["isReal"]] for k, v in c1.items()})
    k2 = OL.pair_keys({k: v[c2["isReal"]] for k, v in c2.items()})
    assert np.array_equal(k1, k2)
    # both lists are sorted by (min id, max id)
    assert np.array_equal(c1["forceLocal"][c1["isReal"]], c2["forceLocal"][c2["isReal"]])
    assert np.array_equal(c1["trsf"][c1["isReal"]], c2["trsf"][c2["isReal"]])
    # energies are block sums in work-list order, which depends on the step of the last collider
    # full run (the restore adds one): equal up to summation order
    e1, e2 = r1.energy(), r2.energy()
    assert np.allclose(e1[0], e2[0], rtol=1e-12, atol=0) and np.array_equal(e1[1], e2[1])
    assert r1.stats()["step"] == r2.stats()["step"]
    assert r1.ctx.time(r1.sid) == r2.ctx.time(r2.sid)


def _scene(kind):
    from woodem_amd._native import PREC_MIXED, Context
    ctx = Context(0)
    if kind == "box":
        a = scenes.box_lattice(nx=8, ny=6, nz=5)
        p = runner.params_from_meta(a, trackEnergy=True)
    elif kind == "periodic":
        a = scenes.relax_and_order(scenes.periodic_hertz(n=4000, seed=4), ctx, iters=100)
        p = runner.params_from_meta(a, trackEnergy=True)
    elif kind == "mixed":
        a = scenes.relax_and_order(scenes.periodic_hertz(n=4000, seed=4), ctx, iters=100)
        p = runner.params_from_meta(a, precision=PREC_MIXED)
    else:
        a = scenes.drum(n=3000, seed=7)
        p = runner.params_from_meta(a)
    return ctx, a, p, runner.cell_from_meta(a)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["box", "periodic", "mixed", "drum"])
def test_checkpoint_continues_bit_for_bit(kind, tmp_path):
    from woodem_amd._native import make_dyndt
    ctx, a, p, cell = _scene(kind)
    r1 = runner.HipRun(a, p, cell, ctx=ctx)
    if kind == "box":
        ctx.set_dyndt(r1.sid, make_dyndt(stepPeriod=50, maxRelInc=1e-3))
    r1.step(2700 if kind == "box" else 300)
    f = snapshot.save(r1, str(tmp_path / ("%s.npz" % kind)))
    r2 = snapshot.load(f, ctx=ctx)
    for n in (1, 37, 400):
        r1.step(n)
        r2.step(n)
        _compare(r1, r2)
    if kind == "box":
        d1, d2 = ctx.dyndt(r1.sid), ctx.dyndt(r2.sid)
        assert (d1.nDone, d1.stepLast, d1.stepPrev) == (d2.nDone, d2.stepLast, d2.stepPrev)
    r1.close()
    r2.close()


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["box", "periodic"])
def test_checkpoint_of_a_restored_scene(kind, tmp_path):
    """save, load, save again before any step, load that: the second checkpoint holds the same
    contact points as the first (they live in the contacts' state rows, not in the gather
    records a ContactLoop writes), and the scene restored from it steps on bit for bit like the
    uninterrupted run (L6Geom's update reads the previous contact point)."""
    ctx, a, p, cell = _scene(kind)
    r1 = runner.HipRun(a, p, cell, ctx=ctx)
    r1.step(2700 if kind == "box" else 300)
    f1 = snapshot.save(r1, str(tmp_path / "first.npz"))
    r2 = snapshot.load(f1, ctx=ctx)
    f2 = snapshot.save(r2, str(tmp_path / "second.npz"))
    r2.close()
    c1, c2 = snapshot.read(f1)[3], snapshot.read(f2)[3]
    assert len(c1["idA"]) > 10
    for k in ("idA", "idB", "cellDist", "trsf", "forceLocal", "phys", "lens", "contPt"):
        assert np.array_equal(c1[k], c2[k]), k
    r3 = snapshot.load(f2, ctx=ctx)
    for n in (1, 37, 400):
        r1.step(n)
        r3.step(n)
        _compare(r1, r3)
    r1.close()
    r3.close()
