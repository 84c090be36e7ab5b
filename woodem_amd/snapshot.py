"""Checkpoints of a scene on the HIP path (npz) and VTK output of its state.

Woo saves whole scenes with boost::serialization (lib/object/ObjectIO.hpp:49-90), a format
that cannot be read without Boost; this module gives the timestep path a neutral one.

Checkpoint schema "woodem_amd.snapshot/1": one .npz archive written by numpy (no pickled
objects; it is read with allow_pickle=False), holding

  meta            JSON text: schema, step, time, dt, periodic, homoDeform, params (the
                  wdm_params fields), dyndt (the wdm_dyndt fields), the arrays' meta
  nodes           pos (N,3), ori (N,4; w,x,y,z), vel, angVel (N,3), mass (N), inertia (N,3),
                  flags (N), force, torque (N,3): DemData after the last step; force and
                  torque are the last ContactLoop's sums with gravity, what the next
                  Leapfrog reads
  particles       shape, nodes (P,3), radius, axis, sense, mask, mat
  materials       density, young, ktDivKn, tanPhi
  clumps          clump_master, clump_off, clump_members, clump_relPos, clump_relOri (when
                  the scene has clumps)
  cell            hSize, gradV, nextGradV (3,3)
  contacts        con_idA, con_idB (functor order), con_cellDist (C,3), con_trsf (C,3,3),
                  con_forceLocal (C,3), con_phys (C,5), con_lens (C,2), con_contPt (C,3):
                  the real contacts (the collider finds the potential ones again)
  energy          energy_values (8), energy_created (8), slots in EnergyTracker order

Restoring a checkpoint and stepping on gives the same trajectory, bit for bit, as stepping
the saved scene on: the first step after a restore runs a collider full run, which keeps
the real contacts with their state and finds the potential ones again; every pair that
touches is among them either way (all nodes move less than verletDist between full runs,
InsertionSortCollider.cpp:340-350), so no force changes.  Energies are block sums in the
order of the contact work list, which depends on the step of the last full run: they agree
up to summation order.

VTK output follows VtkExport (pkg/dem/VtkExport.cpp:132-560) in its `ascii` mode: spheres
as vertices with point data radius, mass, id, mask, vel, angVel, matId
(`{out}spheres.{step}.vtu`); facets as triangles with cell data vel, matId
(`{out}mesh.{step}.vtu`); real contacts as lines between the particles (a sphere's centre,
the contact point for walls and facets) with cell data Fn and |Ft| (`{out}con.{step}.vtp`).  Periodic scenes: sphere points are wrapped into the cell
(Cell::canonicalizePt) and a contact's line starts at the wrapped point of its first
particle.
"""
import base64
import json
import os
import struct

import numpy as np

from . import soa
from ._native import DYNDT_FIELDS, WdmDynDt, WdmParams

SCHEMA = "woodem_amd.snapshot/1"
PARAM_FIELDS = [f for f, _ in WdmParams._fields_]


def _params_dict(p):
    out = {}
    for f in PARAM_FIELDS:
        v = getattr(p, f)
        out[f] = [float(x) for x in v] if f == "gravity" else (float(v) if isinstance(v, float) else int(v))
    return out


def _params_from(d):
    p = WdmParams()
    for f in PARAM_FIELDS:
        if f == "gravity":
            for k in range(3):
                p.gravity[k] = d[f][k]
        else:
            setattr(p, f, d[f])
    return p


def save(run, path):
    """Write the checkpoint of a runner.HipRun (its scene's current state) to `path`."""
    ctx, sid, a = run.ctx, run.sid, run.a
    st = ctx.stats(sid)
    time, dt = ctx.time(sid)
    h, g = ctx.cell(sid)
    nd = ctx.download_nodes(sid, a.n_node)
    c = ctx.contacts(sid)
    phys, lens = ctx.contact_state(sid)
    r = c["isReal"]
    ev, ec = ctx.energy(sid)
    d = ctx.dyndt(sid)
    params = _params_dict(run.params)
    params["dt"] = dt
    meta = dict(schema=SCHEMA, step=int(st["step"]), time=time, dt=dt, periodic=bool(run.cell0["periodic"]),
                homoDeform=int(run.cell0["homoDeform"]), params=params,
                dyndt={f: (float(getattr(d, f)) if isinstance(getattr(d, f), float) else int(getattr(d, f))) for f in DYNDT_FIELDS},
                arrays_meta={k: v for k, v in a.meta.items() if isinstance(v, (int, float, str, bool, list, tuple))})
    arrs = dict(
        meta=np.array(json.dumps(meta)),
        pos=nd["pos"], ori=nd["ori"], vel=nd["vel"], angVel=nd["angVel"], mass=a.mass, inertia=a.inertia, flags=a.flags,
        force=nd["force"], torque=nd["torque"],
        shape=a.shape, nodes=a.nodes, radius=a.radius, axis=a.axis, sense=a.sense, mask=a.mask, mat=a.mat,
        density=a.density, young=a.young, ktDivKn=a.ktDivKn, tanPhi=a.tanPhi,
        hSize=h, gradV=g, nextGradV=np.asarray(run.cell0["nextGradV"], np.float64),
        con_idA=c["idA"][r], con_idB=c["idB"][r], con_cellDist=c["cellDist"][r], con_trsf=c["trsf"][r],
        con_forceLocal=c["forceLocal"][r], con_phys=phys[r], con_lens=lens[r], con_contPt=c["contPt"][r],
        energy_values=ev, energy_created=ec.astype(np.int32),
    )
    if a.clumps:
        for k, v in a.clumps.items():
            arrs["clump_" + k] = np.asarray(v)
    np.savez(path, **arrs)
    return path


def read(path):
    """The checkpoint's contents: (DemArrays at the saved state, meta dict, cell dict,
    contacts dict, (energy values, created))."""
    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(str(z["meta"]))
        if meta.get("schema") != SCHEMA:
            raise ValueError("not a %s checkpoint: %r" % (SCHEMA, meta.get("schema")))
        clumps = None
        if "clump_master" in z:
            clumps = {k: z["clump_" + k] for k in ("master", "off", "members", "relPos", "relOri")}
        a = soa.DemArrays(pos=z["pos"], ori=z["ori"], vel=z["vel"], angVel=z["angVel"], mass=z["mass"], inertia=z["inertia"],
                          flags=z["flags"], force=z["force"], torque=z["torque"], shape=z["shape"], nodes=z["nodes"],
                          radius=z["radius"], axis=z["axis"], sense=z["sense"], mask=z["mask"], mat=z["mat"],
                          density=z["density"], young=z["young"], ktDivKn=z["ktDivKn"], tanPhi=z["tanPhi"],
                          meta=dict(meta["arrays_meta"]), clumps=clumps).normalized()
        cell = dict(periodic=meta["periodic"], hSize=z["hSize"], gradV=z["gradV"], nextGradV=z["nextGradV"],
                    homoDeform=meta["homoDeform"])
        con = {k[4:]: z[k] for k in z.files if k.startswith("con_")}
        energy = (z["energy_values"], z["energy_created"].astype(bool))
    return a, meta, cell, con, energy


def load(path, ctx=None):
    """A runner.HipRun holding the checkpoint's scene, ready to step on."""
    from .runner import HipRun
    a, meta, cell, con, (ev, ec) = read(path)
    run = HipRun(a, _params_from(meta["params"]), cell, ctx=ctx)
    ctx = run.ctx
    ctx.upload_contacts(run.sid, con["idA"], con["idB"], con["cellDist"], con["trsf"], con["forceLocal"], con["phys"],
                        con["lens"], con["contPt"])
    ctx.set_energy(run.sid, ev, ec)
    ctx.set_time(run.sid, meta["step"], meta["time"])
    d = WdmDynDt()
    for f in DYNDT_FIELDS:
        setattr(d, f, meta["dyndt"][f])
    if d.enabled:
        ctx.set_dyndt(run.sid, d)
    return run


# ---------------- VTK (VtkExport, ascii mode) ----------------
def _fmt(x):
    return " ".join("%.17g" % v for v in np.asarray(x, np.float64).ravel())


def _ifmt(x):
    return " ".join("%d" % v for v in np.asarray(x).ravel())


_VTK_DTYPES = {"Float64": "<f8", "Float32": "<f4", "Int32": "<i4", "Int64": "<i8", "UInt8": "u1", "UInt32": "<u4"}


def _arr(name, data, ncomp, kind="Float64", binary=True):
    """A DataArray: VTK's inline binary form (base64 of a UInt32 byte count followed by the
    little-endian values) or ascii"""
    if binary:
        raw = np.ascontiguousarray(np.asarray(data).ravel(), _VTK_DTYPES[kind]).tobytes()
        body = base64.b64encode(struct.pack("<I", len(raw)) + raw).decode("ascii")
        return '<DataArray type="%s" Name="%s" NumberOfComponents="%d" format="binary">%s</DataArray>' % (kind, name, ncomp, body)
    body = _fmt(data) if kind == "Float64" else _ifmt(data)
    return '<DataArray type="%s" Name="%s" NumberOfComponents="%d" format="ascii">%s</DataArray>' % (kind, name, ncomp, body)


def _wrap(pos, h):
    """Cell::canonicalizePt for an unsheared cell: each coordinate into [0, L)"""
    L = np.diag(h)
    return pos - np.floor(pos / L) * L


def _particle_stress(a, c, n_par):
    """DemFuncs::particleStress (pkg/dem/Funcs.cpp:227-240) of every particle: over its real
    contacts, in the order of the other particle's id (Particle::contacts is a map by id),
    sigN += trsf (Fn, 0, 0) / contA and sigT += trsf (0, Ft1, Ft2) / contA; contA is
    L6Geom's contact area, pi rr^2 with rr the smaller sphere radius of a sphere pair and the
    sphere's radius otherwise (L6Geom.cpp:61-63 with the Cg2 functors' r1, r2)."""
    sigN, sigT = np.zeros((n_par, 3)), np.zeros((n_par, 3))
    r = c["isReal"]
    A, B, tr, fl = c["idA"][r], c["idB"][r], c["trsf"][r], c["forceLocal"][r]
    if len(A) == 0:
        return sigN, sigT
    ss = (a.shape[A] == soa.SHAPE_SPHERE) & (a.shape[B] == soa.SHAPE_SPHERE)
    rr = np.where(ss, np.minimum(a.radius[A], a.radius[B]), a.radius[B])
    inv = 1. / (np.pi * (rr * rr))
    fn = np.zeros_like(fl)
    fn[:, 0] = fl[:, 0]
    ft = fl.copy()
    ft[:, 0] = 0.
    gN = np.einsum("kij,kj->ki", tr, fn) * inv[:, None]
    gT = np.einsum("kij,kj->ki", tr, ft) * inv[:, None]
    # each particle's contacts by the other particle's id: entries (particle, other, contact)
    part = np.concatenate([A, B])
    other = np.concatenate([B, A])
    k = np.concatenate([np.arange(len(A)), np.arange(len(A))])
    order = np.lexsort((other, part))
    np.add.at(sigN, part[order], gN[k[order]])
    np.add.at(sigT, part[order], gT[k[order]])
    return sigN, sigT


def vtk_export(run, out, what=("spheres", "mesh", "con"), sphereSphereOnly=False, nanValue=0., binary=True):
    """Write the scene's current state as VTK XML files laid out like VtkExport's
    (pkg/dem/VtkExport.cpp:123-560); returns {'spheres': path, 'mesh': path, 'con': path} for
    the parts written.  Arrays are inline base64 binary (ascii with binary=False).
      spheres: a vertex per sphere (position wrapped into a periodic cell), point data radius,
        mass, id, mask, color, vel, angVel, matId, cell data sigT, sigN;
      mesh: facets as triangles, each with its own points (one triangle; with halfThick, the
        two faces and their closed flat edges as for thickFacetDiv=1), cell data color,
        matState, matId, vel, angVel (of the facet's first node) and |sigma| (0);
      con: the node position of every particle id (0 for a removed particle; wrapped into a
        periodic cell) then added points, a line per real contact between the sphere
        centres (the contact point for a side that is no sphere), a contact crossing the
        periodic boundary drawn twice with a shifted end point; cell data Fn and |Ft|.
    This data model holds no render color and no material state: color is 0, matState
    nanValue.  Walls carry no display extents (Wall.glAB) and are left out (VtkExport with
    infError=False)."""
    def arr(*x, **k):  # every array of this export in the chosen form
        return _arr(*x, binary=binary, **k)

    ctx, sid, a = run.ctx, run.sid, run.a
    step = ctx.stats(sid)["step"]
    h, _ = ctx.cell(sid)
    periodic = bool(run.cell0["periodic"])
    L = np.diag(h)
    nd = ctx.download_nodes(sid, a.n_node)
    pos = nd["pos"]
    files = {}
    d = os.path.dirname(out)
    if d:
        os.makedirs(d, exist_ok=True)
    c = ctx.contacts(sid) if ("con" in what or "spheres" in what) else None
    if "spheres" in what:
        sph = np.nonzero(a.shape == soa.SHAPE_SPHERE)[0]
        nodes = a.nodes[sph, 0]
        spos = pos[nodes]
        if periodic:
            spos = _wrap(spos, h)
        n = len(sph)
        sigN, sigT = _particle_stress(a, c, a.n_par)
        xml = ['<?xml version="1.0"?>', '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">',
               "<UnstructuredGrid>", '<Piece NumberOfPoints="%d" NumberOfCells="%d">' % (n, n),
               "<Points>" + arr("Points", spos, 3) + "</Points>",
               "<Cells>" + arr("connectivity", np.arange(n), 1, "Int32") + arr("offsets", np.arange(1, n + 1), 1, "Int32")
               + arr("types", np.ones(n, np.int32), 1, "UInt8") + "</Cells>",
               "<PointData>", arr("radius", a.radius[sph], 1), arr("mass", a.mass[nodes], 1), arr("id", sph, 1, "Int32"),
               arr("mask", a.mask[sph], 1, "Int32"), arr("color", np.zeros(n), 1), arr("vel", nd["vel"][nodes], 3),
               arr("angVel", nd["angVel"][nodes], 3), arr("matId", a.mat[sph], 1, "Int32"), "</PointData>",
               "<CellData>", arr("sigT", sigT[sph], 3), arr("sigN", sigN[sph], 3), "</CellData>",
               "</Piece>", "</UnstructuredGrid>", "</VTKFile>"]
        fn = "%sspheres.%d.vtu" % (out, step)
        with open(fn, "w") as f:
            f.write("\n".join(xml) + "\n")
        files["spheres"] = fn
    fac = np.nonzero(a.shape == soa.SHAPE_FACET)[0]
    if "mesh" in what and len(fac):
        pts, tris, cellOf = [], [], []
        for p in fac:
            A, B, C = pos[a.nodes[p, 0]], pos[a.nodes[p, 1]], pos[a.nodes[p, 2]]
            i0 = len(pts)
            ht = a.radius[p]
            if ht == 0.:
                pts += [A, B, C]
                t = [(0, 1, 2)]
            else:  # closed flat edges (thickFacetDiv=1, VtkExport.cpp:300-313)
                nn = np.cross(B - A, C - A)
                dz = nn / np.linalg.norm(nn) * ht
                pts += [A + dz, B + dz, C + dz, A - dz, B - dz, C - dz]
                t = [(0, 1, 2), (3, 4, 5), (0, 1, 3), (3, 1, 4), (1, 2, 4), (4, 2, 5), (2, 0, 5), (5, 0, 3)]
            tris += [(i0 + x, i0 + y, i0 + z) for x, y, z in t]
            cellOf += [p] * len(t)
        cellOf = np.array(cellOf)
        nf, nv = len(tris), len(pts)
        n0 = a.nodes[cellOf, 0]
        xml = ['<?xml version="1.0"?>', '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">',
               "<UnstructuredGrid>", '<Piece NumberOfPoints="%d" NumberOfCells="%d">' % (nv, nf),
               "<Points>" + arr("Points", np.array(pts), 3) + "</Points>",
               "<Cells>" + arr("connectivity", np.array(tris), 1, "Int32") + arr("offsets", 3 * np.arange(1, nf + 1), 1, "Int32")
               + arr("types", np.full(nf, 5, np.int32), 1, "UInt8") + "</Cells>",
               "<CellData>", arr("color", np.zeros(nf), 1), arr("matState", np.full(nf, float(nanValue)), 1),
               arr("matId", a.mat[cellOf], 1, "Int32"), arr("vel", nd["vel"][n0], 3), arr("angVel", nd["angVel"][n0], 3),
               arr("|sigma|", np.zeros(nf), 1), "</CellData>",
               "</Piece>", "</UnstructuredGrid>", "</VTKFile>"]
        fn = "%smesh.%d.vtu" % (out, step)
        with open(fn, "w") as f:
            f.write("\n".join(xml) + "\n")
        files["mesh"] = fn
    if "con" in what:
        # the node position of every particle id (VtkExport.cpp:146-160), wrapped with its cell distance
        ppos = np.zeros((a.n_par, 3))
        live = a.shape != 0  # a removed particle has no shape
        ppos[live] = pos[a.nodes[live, 0]]
        wcd = np.zeros((a.n_par, 3), np.int64)
        if periodic:
            wcd = np.floor(ppos / L).astype(np.int64)
            ppos = ppos - wcd * L
            ppos[~live] = 0.
            wcd[~live] = 0
        keep = c["isReal"].copy()
        if sphereSphereOnly:
            keep &= (a.shape[c["idA"]] == soa.SHAPE_SPHERE) & (a.shape[c["idB"]] == soa.SHAPE_SPHERE)
        A, B, cd, cp = c["idA"][keep], c["idB"][keep], c["cellDist"][keep].astype(np.int64), c["contPt"][keep]
        fl = c["forceLocal"][keep]
        nsA, nsB = a.shape[A] != soa.SHAPE_SPHERE, a.shape[B] != soa.SHAPE_SPHERE
        # a contact crossing the periodic boundary is drawn twice (VtkExport.cpp:186-201)
        cross = np.zeros(len(A), bool)
        if periodic:
            cross = ~np.all(cd == wcd[B] - wcd[A], axis=1)
        # points added per contact, in the exporter's order: the contact point for a side that
        # is no sphere (A's, then B's); crossing: A's node shifted next to B's displayed image
        # [, the contact point shifted likewise], then the same for B
        nExtra = nsA.astype(np.int64) + nsB + 2 * cross
        nLines = 1 + cross.astype(np.int64)
        e0 = a.n_par + np.cumsum(nExtra) - nExtra  # first added point and first line of each contact
        l0 = np.cumsum(nLines) - nLines
        extra = np.zeros((int(nExtra.sum()), 3))
        lines = np.zeros((int(nLines.sum()), 2), np.int64)
        ppos_raw = np.zeros((a.n_par, 3))
        ppos_raw[live] = pos[a.nodes[live, 0]]
        q = ~cross
        ia = np.where(nsA, e0, A)
        ib = np.where(nsB, e0 + nsA, B)
        extra[(e0 - a.n_par)[q & nsA]] = cp[q & nsA]
        extra[(e0 + nsA - a.n_par)[q & nsB]] = cp[q & nsB]
        lines[l0[q]] = np.stack([ia[q], ib[q]], 1)
        if cross.any():
            x = cross
            hT = h.T
            f0 = e0[x]
            cA = f0 + 1
            f1 = f0 + 1 + nsA[x]
            cB = f1 + 1
            extra[f0 - a.n_par] = ppos_raw[A[x]] + (wcd[B[x]] - cd[x]).astype(np.float64) @ hT
            extra[f1 - a.n_par] = ppos_raw[B[x]] + (wcd[A[x]] - cd[x]).astype(np.float64) @ hT
            ma, mb = nsA[x], nsB[x]
            extra[cA[ma] - a.n_par] = cp[x][ma] + (wcd[A[x]][ma] - cd[x][ma]).astype(np.float64) @ hT
            extra[cB[mb] - a.n_par] = cp[x][mb] + (wcd[B[x]][mb] - cd[x][mb]).astype(np.float64) @ hT
            lines[l0[x]] = np.stack([np.where(ma, cA, A[x]), f0], 1)
            lines[l0[x] + 1] = np.stack([np.where(mb, cB, B[x]), f1], 1)
        allpts = np.concatenate([ppos, extra])
        m = len(lines)
        conn = lines.reshape(-1)
        Fn = np.nan_to_num(np.repeat(fl[:, 0], nLines), nan=nanValue)
        Ft = np.nan_to_num(np.repeat(np.sqrt(fl[:, 1] ** 2 + fl[:, 2] ** 2), nLines), nan=nanValue)
        xml = ['<?xml version="1.0"?>', '<VTKFile type="PolyData" version="0.1" byte_order="LittleEndian">', "<PolyData>",
               '<Piece NumberOfPoints="%d" NumberOfVerts="0" NumberOfLines="%d" NumberOfStrips="0" NumberOfPolys="0">' % (len(allpts), m),
               "<Points>" + arr("Points", allpts, 3) + "</Points>",
               "<Lines>" + arr("connectivity", conn, 1, "Int32") + arr("offsets", 2 * np.arange(1, m + 1), 1, "Int32")
               + "</Lines>",
               "<CellData>", arr("Fn", Fn, 1), arr("|Ft|", Ft, 1), "</CellData>", "</Piece>", "</PolyData>", "</VTKFile>"]
        fn = "%scon.%d.vtp" % (out, step)
        with open(fn, "w") as f:
            f.write("\n".join(xml) + "\n")
        files["con"] = fn
    return files


def read_vtk_arrays(path):
    """DataArrays of a VTK XML file written by vtk_export, by name (numpy arrays)."""
    import xml.etree.ElementTree as ET
    out = {}
    for e in ET.parse(path).getroot().iter("DataArray"):
        kind = e.get("type")
        if e.get("format") == "binary":
            raw = base64.b64decode((e.text or "").strip())
            nbytes = struct.unpack("<I", raw[:4])[0]
            vals = np.frombuffer(raw[4:4 + nbytes], _VTK_DTYPES[kind])
            vals = vals.astype(np.float64 if kind.startswith("Float") else np.int64)
        else:
            vals = np.array((e.text or "").split(), dtype=np.float64 if kind == "Float64" else np.int64)
        nc = int(e.get("NumberOfComponents", "1"))
        out[e.get("Name")] = vals.reshape(-1, nc) if nc > 1 else vals
    return out
