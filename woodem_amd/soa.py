"""Structure-of-arrays form of a DEM scene: what crosses the C ABI.

Layout rules (see include/woodem_hip.h):
  * uninodal particles (Sphere, Wall) own node index == particle index;
  * facet vertices are extra nodes with index >= n_par;
  * ori is (w, x, y, z); flags are DemData.flags bits (pkg/dem/Particle.hpp:177-182).
"""
from dataclasses import dataclass, field

import numpy as np

from ._native import SHAPE_FACET, SHAPE_SPHERE, SHAPE_WALL

# DemData flag bits, pkg/dem/Particle.hpp:177-182
DOF_X, DOF_Y, DOF_Z, DOF_RX, DOF_RY, DOF_RZ = 1, 2, 4, 8, 16, 32
DOF_ALL = 63
CLUMP_CLUMPED, CLUMP_CLUMP = 64, 128
ENERGY_SKIP, GRAVITY_SKIP, TRACER_SKIP, DAMPING_SKIP = 256, 512, 1024, 2048

# DemField default masks, pkg/dem/Particle.hpp:302-306
DEFAULT_MOVABLE_MASK = 0b0101
DEFAULT_BOUNDARY_MASK = 0b0011
DEFAULT_LONE_MASK = 0b0010


@dataclass
class DemArrays:
    pos: np.ndarray
    ori: np.ndarray
    vel: np.ndarray
    angVel: np.ndarray
    mass: np.ndarray
    inertia: np.ndarray
    flags: np.ndarray
    force: np.ndarray
    torque: np.ndarray
    shape: np.ndarray
    nodes: np.ndarray
    radius: np.ndarray
    axis: np.ndarray
    sense: np.ndarray
    mask: np.ndarray
    mat: np.ndarray
    density: np.ndarray
    young: np.ndarray
    ktDivKn: np.ndarray
    tanPhi: np.ndarray
    meta: dict = field(default_factory=dict)
    # ClumpData (pkg/dem/Clump.hpp): dict(master[nc], off[nc+1], members[m], relPos[m,3], relOri[m,4]) or None
    clumps: dict = None

    @property
    def n_node(self):
        return len(self.pos)

    @property
    def n_par(self):
        return len(self.shape)

    @property
    def n_mat(self):
        return len(self.density)

    def normalized(self):
        """Copy with every array contiguous and of the ABI dtype."""
        f64 = lambda a, s: np.ascontiguousarray(a, dtype=np.float64).reshape(s)
        i32 = lambda a, s: np.ascontiguousarray(a, dtype=np.int32).reshape(s)
        u32 = lambda a, s: np.ascontiguousarray(a, dtype=np.uint32).reshape(s)
        n, p, m = len(self.pos), len(self.shape), len(self.density)
        return DemArrays(
            pos=f64(self.pos, (n, 3)), ori=f64(self.ori, (n, 4)), vel=f64(self.vel, (n, 3)), angVel=f64(self.angVel, (n, 3)),
            mass=f64(self.mass, (n,)), inertia=f64(self.inertia, (n, 3)), flags=u32(self.flags, (n,)),
            force=f64(self.force, (n, 3)), torque=f64(self.torque, (n, 3)),
            shape=i32(self.shape, (p,)), nodes=i32(self.nodes, (p, 3)), radius=f64(self.radius, (p,)),
            axis=i32(self.axis, (p,)), sense=i32(self.sense, (p,)), mask=u32(self.mask, (p,)), mat=i32(self.mat, (p,)),
            density=f64(self.density, (m,)), young=f64(self.young, (m,)), ktDivKn=f64(self.ktDivKn, (m,)),
            tanPhi=f64(self.tanPhi, (m,)), meta=dict(self.meta), clumps=self._norm_clumps())

    def _norm_clumps(self):
        c = self.clumps
        if not c:
            return None
        nm = len(c["members"])
        return dict(master=np.ascontiguousarray(c["master"], np.int32), off=np.ascontiguousarray(c["off"], np.int32),
                    members=np.ascontiguousarray(c["members"], np.int32),
                    relPos=np.ascontiguousarray(c["relPos"], np.float64).reshape(nm, 3),
                    relOri=np.ascontiguousarray(c["relOri"], np.float64).reshape(nm, 4))

    def copy(self):
        return self.normalized()

    def validate(self):
        a = self
        for p in range(a.n_par):
            pass
        uni = (a.shape == SHAPE_SPHERE) | (a.shape == SHAPE_WALL)
        if np.any(a.nodes[uni, 0] != np.nonzero(uni)[0]):
            raise ValueError("uninodal particles must use node index == particle index")
        fac = a.shape == SHAPE_FACET
        if np.any(a.nodes[fac] < a.n_par):
            raise ValueError("facet nodes must have index >= n_par")


def sphere_mass_inertia(radius, density):
    """Sphere::lumpMassInertia, pkg/dem/Sphere.cpp:24-31: m = 4/3 pi r^3 rho, I = 2/5 m r^2."""
    r = np.asarray(radius, dtype=np.float64)
    m = (4 / 3.) * np.pi * (r * r * r) * density
    inertia = (2. / 5.) * m * (r * r)
    return m, inertia


def spheres_arrays(pos, radius, density=1e3, young=1e7, ktDivKn=.2, tanPhi=.5, vel=None, mat=None, materials=None,
                   mask=DEFAULT_MOVABLE_MASK):
    """Arrays for a set of free spheres with one (or a given table of) FrictMat."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    n = len(pos)
    radius = np.broadcast_to(np.asarray(radius, dtype=np.float64), (n,)).copy()
    if materials is None:
        materials = dict(density=[density], young=[young], ktDivKn=[ktDivKn], tanPhi=[tanPhi])
    matIx = np.zeros(n, np.int32) if mat is None else np.asarray(mat, np.int32)
    dens = np.asarray(materials["density"], np.float64)[matIx]
    m, I = sphere_mass_inertia(radius, dens)
    ori = np.zeros((n, 4))
    ori[:, 0] = 1
    return DemArrays(
        pos=pos, ori=ori, vel=np.zeros((n, 3)) if vel is None else np.asarray(vel, np.float64).reshape(n, 3),
        angVel=np.zeros((n, 3)), mass=m, inertia=np.repeat(I[:, None], 3, axis=1), flags=np.zeros(n, np.uint32),
        force=np.zeros((n, 3)), torque=np.zeros((n, 3)),
        shape=np.full(n, SHAPE_SPHERE, np.int32), nodes=np.stack([np.arange(n), -np.ones(n), -np.ones(n)], 1).astype(np.int32),
        radius=radius, axis=np.zeros(n, np.int32), sense=np.zeros(n, np.int32), mask=np.full(n, mask, np.uint32), mat=matIx,
        density=np.asarray(materials["density"], np.float64), young=np.asarray(materials["young"], np.float64),
        ktDivKn=np.asarray(materials["ktDivKn"], np.float64), tanPhi=np.asarray(materials["tanPhi"], np.float64),
    ).normalized()


def append_walls(a, walls, mat=0, mask=DEFAULT_BOUNDARY_MASK):
    """Append axis-aligned walls [(coord, axis, sense), ...] as fixed, massless particles
    (utils.wall, py/utils.py:177-199).  Walls must come after all facets' extra nodes are
    not yet present, i.e. call this before adding facets."""
    a = a.normalized()
    if np.any(a.shape == SHAPE_FACET):
        raise ValueError("append walls before facets")
    k = len(walls)
    n0 = a.n_par
    pos = np.zeros((k, 3))
    ax = np.zeros(k, np.int32)
    se = np.zeros(k, np.int32)
    for i, (coord, axis, sense) in enumerate(walls):
        pos[i, axis] = coord
        ax[i] = axis
        se[i] = sense
    ori = np.zeros((k, 4))
    ori[:, 0] = 1
    cat = lambda x, y: np.concatenate([x, y], 0)
    return DemArrays(
        pos=cat(a.pos, pos), ori=cat(a.ori, ori), vel=cat(a.vel, np.zeros((k, 3))), angVel=cat(a.angVel, np.zeros((k, 3))),
        mass=cat(a.mass, np.zeros(k)), inertia=cat(a.inertia, np.zeros((k, 3))),
        flags=cat(a.flags, np.full(k, DOF_ALL, np.uint32)), force=cat(a.force, np.zeros((k, 3))),
        torque=cat(a.torque, np.zeros((k, 3))), shape=cat(a.shape, np.full(k, SHAPE_WALL, np.int32)),
        nodes=cat(a.nodes, np.stack([np.arange(n0, n0 + k), -np.ones(k), -np.ones(k)], 1).astype(np.int32)),
        radius=cat(a.radius, np.zeros(k)), axis=cat(a.axis, ax), sense=cat(a.sense, se),
        mask=cat(a.mask, np.full(k, mask, np.uint32)), mat=cat(a.mat, np.full(k, mat, np.int32)),
        density=a.density, young=a.young, ktDivKn=a.ktDivKn, tanPhi=a.tanPhi, meta=dict(a.meta)).normalized()


def with_spare(a, k):
    """Arrays with `k` spare particle slots (shape NONE) after the particles, for particles
    inserted later (Context.insert_spheres): each slot has its node (index == slot), fully
    blocked, massless and skipped by the energy terms; facet nodes move up by k."""
    a = a.normalized()
    n0, m = a.n_par, a.n_node
    inert = DOF_ALL | ENERGY_SKIP | GRAVITY_SKIP | DAMPING_SKIP
    ins = lambda x, fill: np.concatenate([x[:n0], fill, x[n0:]], 0)
    z3 = np.zeros((k, 3))
    ori = np.zeros((k, 4))
    ori[:, 0] = 1
    nodes = a.nodes.copy()
    nodes[(a.shape == SHAPE_FACET)[:, None] & (nodes >= n0)] += k
    out = DemArrays(
        pos=ins(a.pos, z3), ori=ins(a.ori, ori), vel=ins(a.vel, z3), angVel=ins(a.angVel, z3), mass=ins(a.mass, np.zeros(k)),
        inertia=ins(a.inertia, z3), flags=ins(a.flags, np.full(k, inert, np.uint32)), force=ins(a.force, z3),
        torque=ins(a.torque, z3), shape=np.concatenate([a.shape, np.zeros(k, np.int32)]),
        nodes=np.concatenate([nodes, -np.ones((k, 3), np.int32)]), radius=np.concatenate([a.radius, np.zeros(k)]),
        axis=np.concatenate([a.axis, np.zeros(k, np.int32)]), sense=np.concatenate([a.sense, np.zeros(k, np.int32)]),
        mask=np.concatenate([a.mask, np.zeros(k, np.uint32)]), mat=np.concatenate([a.mat, np.zeros(k, np.int32)]),
        density=a.density, young=a.young, ktDivKn=a.ktDivKn, tanPhi=a.tanPhi, meta=dict(a.meta))
    if a.clumps:
        c = dict(a.clumps)
        bump = lambda x: np.where(np.asarray(x) >= n0, np.asarray(x) + k, np.asarray(x))
        c["master"], c["members"] = bump(c["master"]), bump(c["members"])
        out.clumps = c
    return out.normalized()


def permute_particles(a, perm):
    """Relabel uninodal particles (no facets) with a new order: new id i = old id perm[i]."""
    a = a.normalized()
    if np.any(a.shape == SHAPE_FACET):
        raise ValueError("permute_particles: facets are not supported")
    perm = np.asarray(perm)
    out = DemArrays(pos=a.pos[perm], ori=a.ori[perm], vel=a.vel[perm], angVel=a.angVel[perm], mass=a.mass[perm],
                    inertia=a.inertia[perm], flags=a.flags[perm], force=a.force[perm], torque=a.torque[perm],
                    shape=a.shape[perm], nodes=a.nodes[perm], radius=a.radius[perm], axis=a.axis[perm],
                    sense=a.sense[perm], mask=a.mask[perm], mat=a.mat[perm], density=a.density, young=a.young,
                    ktDivKn=a.ktDivKn, tanPhi=a.tanPhi, meta=dict(a.meta))
    out.nodes[:, 0] = np.arange(len(perm))
    return out.normalized()
