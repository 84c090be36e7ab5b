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


def vtk_export(run, out, what=("spheres", "mesh", "con"), sphereSphereOnly=False, nanValue=0., binary=True):
    """Write the scene's current state as VTK XML files named like VtkExport's; returns
    {'spheres': path, 'mesh': path, 'con': path} for the parts written.  Arrays are inline
    base64 binary (ascii with binary=False)."""
    def arr(*x, **k):  # every array of this export in the chosen form
        return _arr(*x, binary=binary, **k)

    ctx, sid, a = run.ctx, run.sid, run.a
    step = ctx.stats(sid)["step"]
    h, _ = ctx.cell(sid)
    periodic = bool(run.cell0["periodic"])
    nd = ctx.download_nodes(sid, a.n_node)
    pos = nd["pos"]
    files = {}
    d = os.path.dirname(out)
    if d:
        os.makedirs(d, exist_ok=True)
    sph = np.nonzero(a.shape == soa.SHAPE_SPHERE)[0]
    spos = pos[a.nodes[sph, 0]]
    if periodic:
        spos = _wrap(spos, h)
    if "spheres" in what:
        n = len(sph)
        nodes = a.nodes[sph, 0]
        xml = ['<?xml version="1.0"?>', '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">',
               "<UnstructuredGrid>", '<Piece NumberOfPoints="%d" NumberOfCells="%d">' % (n, n),
               "<Points>" + arr("Points", spos, 3) + "</Points>",
               "<Cells>" + arr("connectivity", np.arange(n), 1, "Int32") + arr("offsets", np.arange(1, n + 1), 1, "Int32")
               + arr("types", np.ones(n, np.int32), 1, "UInt8") + "</Cells>",
               "<PointData>", arr("radius", a.radius[sph], 1), arr("mass", a.mass[nodes], 1), arr("id", sph, 1, "Int32"),
               arr("mask", a.mask[sph], 1, "Int32"), arr("vel", nd["vel"][nodes], 3), arr("angVel", nd["angVel"][nodes], 3),
               arr("matId", a.mat[sph], 1, "Int32"), "</PointData>", "</Piece>", "</UnstructuredGrid>", "</VTKFile>"]
        fn = "%sspheres.%d.vtu" % (out, step)
        with open(fn, "w") as f:
            f.write("\n".join(xml) + "\n")
        files["spheres"] = fn
    fac = np.nonzero(a.shape == soa.SHAPE_FACET)[0]
    if "mesh" in what and len(fac):
        vn = np.unique(a.nodes[fac].ravel())
        local = {int(q): i for i, q in enumerate(vn)}
        conn = np.array([[local[int(q)] for q in a.nodes[p]] for p in fac])
        vel = nd["vel"][a.nodes[fac]].mean(axis=1)
        nf = len(fac)
        xml = ['<?xml version="1.0"?>', '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">',
               "<UnstructuredGrid>", '<Piece NumberOfPoints="%d" NumberOfCells="%d">' % (len(vn), nf),
               "<Points>" + arr("Points", pos[vn], 3) + "</Points>",
               "<Cells>" + arr("connectivity", conn, 1, "Int32") + arr("offsets", 3 * np.arange(1, nf + 1), 1, "Int32")
               + arr("types", np.full(nf, 5, np.int32), 1, "UInt8") + "</Cells>",
               "<CellData>", arr("vel", vel, 3), arr("matId", a.mat[fac], 1, "Int32"), "</CellData>",
               "</Piece>", "</UnstructuredGrid>", "</VTKFile>"]
        fn = "%smesh.%d.vtu" % (out, step)
        with open(fn, "w") as f:
            f.write("\n".join(xml) + "\n")
        files["mesh"] = fn
    if "con" in what:
        c = ctx.contacts(sid)
        keep = c["isReal"].copy()
        if sphereSphereOnly:
            keep &= (a.shape[c["idA"]] == soa.SHAPE_SPHERE) & (a.shape[c["idB"]] == soa.SHAPE_SPHERE)
        A, B, cd = c["idA"][keep], c["idB"][keep], c["cellDist"][keep]
        cp = c["contPt"][keep]
        # a sphere's end is its centre, any other particle's the contact point (VtkExport.cpp:180-190)
        sa = (a.shape[A] == soa.SHAPE_SPHERE)[:, None]
        sb = (a.shape[B] == soa.SHAPE_SPHERE)[:, None]
        pa = np.where(sa, pos[a.nodes[A, 0]], cp)
        pb = np.where(sb, pos[a.nodes[B, 0]] + cd @ h.T, cp)
        if periodic:
            start = _wrap(pa, h)
            ends = start + (pb - pa)
        else:
            start, ends = pa, pb
        m = len(A)
        pts = np.empty((2 * m, 3))
        pts[0::2], pts[1::2] = start, ends
        fl = c["forceLocal"][keep]
        Fn = np.nan_to_num(fl[:, 0], nan=nanValue)
        Ft = np.nan_to_num(np.sqrt(fl[:, 1] ** 2 + fl[:, 2] ** 2), nan=nanValue)
        xml = ['<?xml version="1.0"?>', '<VTKFile type="PolyData" version="0.1" byte_order="LittleEndian">', "<PolyData>",
               '<Piece NumberOfPoints="%d" NumberOfVerts="0" NumberOfLines="%d" NumberOfStrips="0" NumberOfPolys="0">' % (2 * m, m),
               "<Points>" + arr("Points", pts, 3) + "</Points>",
               "<Lines>" + arr("connectivity", np.arange(2 * m), 1, "Int32") + arr("offsets", 2 * np.arange(1, m + 1), 1, "Int32")
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
