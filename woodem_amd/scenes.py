"""Seeded generators for the benchmark scenes (as DemArrays + engine settings).

box_lattice: 10,000 monodisperse spheres in a 6-wall box under gravity
  (25x20x20 lattice, 11 mm spacing, +-0.4 mm jitter, r=5 mm, rho=3000,
  FrictMat E=1e7, ktDivKn=.2, tanPhi=.5, IdealElPl, dt=0.3 p-wave, damping .2).
  Placement (not fixed by the workload text): lattice centred in x and y, first
  layer 6 mm above the floor (1 mm gap); box (0,0,0)-(0.3,0.25,0.25).
periodic_hertz: N spheres, lognormal radii (median 1 mm, sigma .3, clipped to
  [0.5, 3] mm) in a periodic cube at 0.60 solid fraction, Hertz E=7e10, nu=.3,
  en=.7, tanPhi=.4, Gaussian velocities (std 0.05 m/s), no gravity, cell strain
  rate 1e-3/s (isotropic).  Not fixed by the workload text and chosen here:
  density 2650 kg/m3, dt = 0.3 r_min / sqrt(E/rho); the overlaps of the uniform
  placement are relaxed on the device (wdm_relax_overlaps); particle ids are
  then renumbered along a Morton curve so that neighbours have nearby ids.
deposition_box: one scene of the 512-scene batch workload: N = round(exp(U(ln 500,
  ln 2e4))) spheres, r ~ U[2, 6] mm, tanPhi ~ U[.1, .8], Leapfrog damping ~ U[.05, .4],
  all drawn from the scene's own seed, in a closed cubic box (6 walls) whose volume is
  the solid volume / 0.40.  Not fixed by the workload text and chosen here: the cube
  shape; centres uniform in the box, overlaps then relaxed on the device with the walls
  holding the spheres in (relax_box); FrictMat E=1e7, rho=3000, ktDivKn=.2, IdealElPl,
  gravity -9.81 z, dt = 0.3 r_min / sqrt(E/rho).
"""
import math

import numpy as np

from . import soa as _soa


def box_lattice(nx=25, ny=20, nz=20, seed=1, spacing=11e-3, jitter=0.4e-3, radius=5e-3, density=3000., young=1e7,
                ktDivKn=.2, tanPhi=.5, box=((0., 0., 0.), (.3, .25, .25)), gap=1e-3):
    rng = np.random.default_rng(seed)
    mn, mx = np.asarray(box[0]), np.asarray(box[1])
    span = np.array([(nx - 1) * spacing, (ny - 1) * spacing, (nz - 1) * spacing])
    origin = np.array([mn[0] + .5 * (mx[0] - mn[0] - span[0]), mn[1] + .5 * (mx[1] - mn[1] - span[1]), mn[2] + radius + gap])
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    pos = origin + spacing * np.stack([ii.ravel(), jj.ravel(), kk.ravel()], 1).astype(np.float64)
    pos = pos + rng.uniform(-jitter, jitter, size=pos.shape)
    a = _soa.spheres_arrays(pos, radius, density=density, young=young, ktDivKn=ktDivKn, tanPhi=tanPhi)
    walls = []
    for sense, coord in ((1, mn), (-1, mx)):
        for ax in (0, 1, 2):
            walls.append((float(coord[ax]), ax, sense))
    a = _soa.append_walls(a, walls)
    dt = 0.3 * radius / math.sqrt(young / density)
    a.meta.update(dict(kind="box_lattice", dt=dt, gravity=(0., 0., -9.81), damping=.2, law="IdealElPl", periodic=False,
                       n_spheres=nx * ny * nz, seed=seed))
    return a


def lognormal_radii(n, rng, median=1e-3, sigma=.3, lo=.5e-3, hi=3e-3):
    r = np.exp(math.log(median) + sigma * rng.standard_normal(n))
    return np.clip(r, lo, hi)


def morton_order(pos, L):
    """Permutation sorting points along a 3-D Morton (Z-order) curve of 2^10 cells per axis."""
    q = np.floor((np.mod(pos, L) / L) * 1024).astype(np.int64)
    q = np.clip(q, 0, 1023)

    def spread(v):
        v = v & 0x3FF
        v = (v | (v << 16)) & 0x30000FF
        v = (v | (v << 8)) & 0x300F00F
        v = (v | (v << 4)) & 0x30C30C3
        v = (v | (v << 2)) & 0x9249249
        return v

    code = spread(q[:, 0]) | (spread(q[:, 1]) << 1) | (spread(q[:, 2]) << 2)
    return np.argsort(code, kind="stable")


def periodic_hertz(n=1_000_000, seed=2, phi=.60, density=2650., young=7e10, poisson=.3, en=.7, tanPhi=.4, velStd=.05,
                   strainRate=1e-3):
    """Initial state before relaxation (positions uniform, overlapping)."""
    rng = np.random.default_rng(seed)
    r = lognormal_radii(n, rng)
    vol = float(np.sum((4. / 3.) * math.pi * r ** 3))
    L = (vol / phi) ** (1. / 3.)
    pos = rng.uniform(0, L, size=(n, 3))
    vel = rng.normal(0, velStd, size=(n, 3))
    a = _soa.spheres_arrays(pos, r, density=density, young=young, ktDivKn=.2, tanPhi=tanPhi, vel=vel)
    dt = 0.3 * float(r.min()) / math.sqrt(young / density)
    a.meta.update(dict(kind="periodic_hertz", dt=dt, gravity=(0., 0., 0.), damping=0., law="Hertz", poisson=poisson, en=en,
                       periodic=True, L=L, strainRate=strainRate, n_spheres=n, seed=seed, phi=phi, density=density))
    return a


def relax_and_order(a, ctx, iters=500):
    """Relax overlaps on the device, then renumber along a Morton curve."""
    sid = ctx.scene_alloc(a.n_par, a.n_node, a.n_mat)
    try:
        ctx.upload(sid, a)
        L = a.meta["L"]
        ctx.set_cell(sid, True, np.diag([L, L, L]), np.zeros((3, 3)), np.zeros((3, 3)), 4)
        ctx.relax(sid, iters)
        st = ctx.download_nodes(sid, a.n_node)
    finally:
        ctx.scene_free(sid)
    b = a.copy()
    b.pos = np.mod(st["pos"], a.meta["L"])
    b = _soa.permute_particles(b, morton_order(b.pos, a.meta["L"]))
    b.meta["relaxIters"] = iters
    return b


def hertz_params_for(a):
    """Engine attributes of the periodic Hertz workload, as a dict for woodem_amd.bind."""
    return dict(dt=a.meta["dt"], damping=0., gravity=(0., 0., 0.), law=1, poisson=a.meta["poisson"], en=a.meta["en"])


def facet_floor(n=6, tilt=.2, cells=6, size=.1, radius=4e-3, spacing=9e-3, seed=3, density=2000., young=1e7, tanPhi=.5,
                halfThick=0.):
    """Spheres (n x n x 2 lattice, jittered) falling on a tilted square floor made of
    2*cells^2 fixed facets sharing vertex nodes (utils.facet, py/utils.py:242-268)."""
    rng = np.random.default_rng(seed)
    ii, jj, kk = np.meshgrid(np.arange(n), np.arange(n), np.arange(2), indexing="ij")
    pos = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], 1).astype(np.float64) * spacing
    pos += np.array([.5 * (size - (n - 1) * spacing), .5 * (size - (n - 1) * spacing), .022])
    pos += rng.uniform(-.5e-3, .5e-3, pos.shape)
    a = _soa.spheres_arrays(pos, radius, density=density, young=young, tanPhi=tanPhi)
    ns = a.n_par
    # vertex grid, tilted about the x axis
    g = np.linspace(0, size, cells + 1)
    X, Y = np.meshgrid(g, g, indexing="ij")
    Z = Y * math.tan(tilt)
    verts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    vid = lambda i, j: i * (cells + 1) + j
    tris = []
    for i in range(cells):
        for j in range(cells):
            tris.append((vid(i, j), vid(i + 1, j), vid(i + 1, j + 1)))
            tris.append((vid(i, j), vid(i + 1, j + 1), vid(i, j + 1)))
    nf = len(tris)
    npar = ns + nf
    nv = len(verts)
    nn = npar + nv
    ori = np.zeros((nn, 4))
    ori[:, 0] = 1
    pos_n = np.concatenate([a.pos, np.zeros((nf, 3)), verts])
    flags = np.concatenate([a.flags, np.full(nf + nv, _soa.DOF_ALL, np.uint32)])
    nodes = np.concatenate([a.nodes, np.array(tris, np.int32) + npar])
    b = _soa.DemArrays(
        pos=pos_n, ori=ori, vel=np.zeros((nn, 3)), angVel=np.zeros((nn, 3)),
        mass=np.concatenate([a.mass, np.zeros(nf + nv)]), inertia=np.concatenate([a.inertia, np.zeros((nf + nv, 3))]),
        flags=flags, force=np.zeros((nn, 3)), torque=np.zeros((nn, 3)),
        shape=np.concatenate([a.shape, np.full(nf, 3, np.int32)]), nodes=nodes,
        radius=np.concatenate([a.radius, np.full(nf, halfThick)]), axis=np.zeros(npar, np.int32),
        sense=np.zeros(npar, np.int32), mask=np.concatenate([a.mask, np.full(nf, _soa.DEFAULT_BOUNDARY_MASK, np.uint32)]),
        mat=np.zeros(npar, np.int32), density=a.density, young=a.young, ktDivKn=a.ktDivKn, tanPhi=a.tanPhi).normalized()
    b.validate()
    dt = 0.3 * radius / math.sqrt(young / density)
    b.meta.update(dict(kind="facet_floor", dt=dt, gravity=(0., 0., -9.81), damping=.2, law="IdealElPl", periodic=False,
                       n_spheres=ns, n_facets=nf, seed=seed))
    return b


def deposition_box(seed=0, n=None, rmin=2e-3, rmax=6e-3, phi=.40, density=3000., young=1e7, ktDivKn=.2, nmin=500,
                   nmax=20000):
    """Initial state before relaxation (uniform centres, overlapping); see relax_box."""
    rng = np.random.default_rng(seed)
    N = int(n) if n is not None else int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
    r = rng.uniform(rmin, rmax, N)
    tanPhi = float(rng.uniform(.1, .8))
    damping = float(rng.uniform(.05, .4))
    side = (float(np.sum((4. / 3.) * math.pi * r ** 3)) / phi) ** (1. / 3.)
    pos = r[:, None] + rng.uniform(0, 1, (N, 3)) * (side - 2 * r)[:, None]
    a = _soa.spheres_arrays(pos, r, density=density, young=young, ktDivKn=ktDivKn, tanPhi=tanPhi)
    walls = [(0., 0, 1), (side, 0, -1), (0., 1, 1), (side, 1, -1), (0., 2, 1), (side, 2, -1)]
    a = _soa.append_walls(a, walls)
    dt = 0.3 * float(r.min()) / math.sqrt(young / density)
    a.meta.update(dict(kind="deposition_box", dt=dt, gravity=(0., 0., -9.81), damping=damping, law="IdealElPl",
                       periodic=False, n_spheres=N, seed=seed, side=side, tanPhi=tanPhi, phi=phi))
    return a


def relax_box(a, ctx, iters=300):
    """Relax the overlaps of a walled scene on the device (walls keep the spheres in)."""
    sid = ctx.scene_alloc(a.n_par, a.n_node, a.n_mat)
    try:
        ctx.upload(sid, a)
        ctx.set_cell(sid, False, np.eye(3), np.zeros((3, 3)), np.zeros((3, 3)), 4)
        ctx.relax(sid, iters)
        st = ctx.download_nodes(sid, a.n_node)
    finally:
        ctx.scene_free(sid)
    b = a.copy()
    b.pos = st["pos"]
    b.meta["relaxIters"] = iters
    return b


def drum(seed=0, n=None, radius=.25, length=.1, segments=64, rings=8, rmin=1e-3, rmax=4e-3, omega=None, nmin=1000,
         nmax=200000, density=3000., young=1e7, ktDivKn=.2, tanPhi=.5, damping=.2, gap=.5e-3):
    """Rotating drum of the serving workload (the pattern of examples/mill.py:28-57): a
    cylinder of `segments` x `rings` quads (2 facets each) around the x axis with two end
    walls, all clumped to a central node at the origin (ClumpData::makeClump of massless
    nodes: every DOF of the master blocked, pkg/dem/Clump.cpp:228-237) that spins at
    omega ~ U[.5, 3] rad/s about x.  Spheres: count round(exp(U(ln 1e3, ln 2e5))) capped at
    the drum's lattice capacity, r ~ U[1, 4] mm, on a cubic lattice of pitch 2 r_max + gap
    inside the drum, lowest sites first.  Not fixed by the workload text and chosen here:
    the lattice fill and its capacity cap (2e5 spheres of this size cannot fit the drum),
    FrictMat E=1e7, rho=3000, ktDivKn=.2, tanPhi=.5, damping .2, IdealElPl,
    dt = 0.3 r_min / sqrt(E/rho)."""
    rng = np.random.default_rng(seed)
    want = int(n) if n is not None else int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
    w = float(omega) if omega is not None else float(rng.uniform(.5, 3.))
    pitch = 2 * rmax + gap
    rin = radius * math.cos(math.pi / segments) - rmax - gap  # inside the flat facets
    xs = np.arange(-.5 * length + rmax + gap, .5 * length - rmax - gap + 1e-12, pitch)
    ys = np.arange(-rin, rin + 1e-12, pitch)
    X, Y, Z = np.meshgrid(xs, ys, ys, indexing="ij")
    sites = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    sites = sites[sites[:, 1] ** 2 + sites[:, 2] ** 2 <= rin * rin]
    sites = sites[np.lexsort((sites[:, 0], sites[:, 1], sites[:, 2]))]  # lowest first
    capacity = len(sites)
    N = min(want, capacity)
    r = rng.uniform(rmin, rmax, N)
    a = _soa.spheres_arrays(sites[:N], r, density=density, young=young, ktDivKn=ktDivKn, tanPhi=tanPhi)
    a = _soa.append_walls(a, [(-.5 * length, 0, 1), (.5 * length, 0, -1)])
    nw = 2
    # facet vertices: rings+1 circles of `segments` points
    th = 2 * math.pi * np.arange(segments) / segments
    xv = np.linspace(-.5 * length, .5 * length, rings + 1)
    verts = np.array([(x, radius * math.cos(t), radius * math.sin(t)) for x in xv for t in th])
    vid = lambda j, i: j * segments + (i % segments)
    tris = []
    for j in range(rings):
        for i in range(segments):
            tris.append((vid(j, i), vid(j + 1, i), vid(j + 1, i + 1)))
            tris.append((vid(j, i), vid(j + 1, i + 1), vid(j, i + 1)))
    nf = len(tris)
    npar0 = a.n_par  # spheres + walls
    npar = npar0 + nf
    nv = len(verts)
    master = npar + nv  # central node
    nn = master + 1
    ori = np.zeros((nn, 4))
    ori[:, 0] = 1
    pos_n = np.concatenate([a.pos, np.zeros((nf, 3)), verts, np.zeros((1, 3))])
    flags = np.concatenate([a.flags, np.full(nf + nv + 1, _soa.DOF_ALL, np.uint32)])
    members = np.concatenate([np.arange(N, N + nw), np.arange(npar, npar + nv)]).astype(np.int32)
    flags[members] |= _soa.CLUMP_CLUMPED
    flags[master] |= _soa.CLUMP_CLUMP
    angVel = np.zeros((nn, 3))
    angVel[master] = (w, 0., 0.)
    # members move with the master from the start (ClumpData::applyToMembers)
    for mi in members:
        angVel[mi] = (w, 0., 0.)
    vel = np.zeros((nn, 3))
    vel[members] = np.cross(angVel[members], pos_n[members])
    b = _soa.DemArrays(
        pos=pos_n, ori=ori, vel=vel, angVel=angVel,
        mass=np.concatenate([a.mass, np.zeros(nf + nv + 1)]), inertia=np.concatenate([a.inertia, np.zeros((nf + nv + 1, 3))]),
        flags=flags, force=np.zeros((nn, 3)), torque=np.zeros((nn, 3)),
        shape=np.concatenate([a.shape, np.full(nf, 3, np.int32)]),
        nodes=np.concatenate([a.nodes, np.array(tris, np.int32) + npar]),
        radius=np.concatenate([a.radius, np.zeros(nf)]), axis=np.concatenate([a.axis, np.zeros(nf, np.int32)]),
        sense=np.concatenate([a.sense, np.zeros(nf, np.int32)]),
        mask=np.concatenate([a.mask, np.full(nf, _soa.DEFAULT_BOUNDARY_MASK, np.uint32)]),
        mat=np.zeros(npar, np.int32), density=a.density, young=a.young, ktDivKn=a.ktDivKn, tanPhi=a.tanPhi,
        clumps=dict(master=np.array([master], np.int32), off=np.array([0, len(members)], np.int32), members=members,
                    relPos=pos_n[members], relOri=ori[members])).normalized()
    b.validate()
    dt = 0.3 * float(r.min()) / math.sqrt(young / density)
    b.meta.update(dict(kind="drum", dt=dt, gravity=(0., 0., -9.81), damping=damping, law="IdealElPl", periodic=False,
                       n_spheres=N, n_requested=want, capacity=capacity, omega=w, n_facets=nf, seed=seed))
    return b
