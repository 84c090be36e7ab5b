"""Array-level driver of one scene on the HIP path (used by bench.py and the tests;
woodem_amd.core.Scene is the object-level, Woo-style front end)."""
import math

import numpy as np

from ._native import LAW_HERTZ_DMT, LAW_IDEALELPL, PREC_FP64, PREC_MIXED, Context, WdmParams


def make_params(dt, damping=0., gravity=(0., 0., 0.), verletDist=-.05, distFactor=1., law=LAW_IDEALELPL, poisson=.2,
                en=float("nan"), noSlip=False, noBreak=False, noFrict=False, noAttraction=True, trackEnergy=False,
                trsfRenorm=100, dist00=True, loneMask=0b0010, precision=PREC_FP64):
    p = WdmParams()
    p.dt, p.damping = dt, damping
    for k in range(3):
        p.gravity[k] = gravity[k]
    p.verletDist, p.distFactor = verletDist, distFactor
    p.poisson, p.en, p.gamma, p.alpha = poisson, en, 0., 0.
    p.law = law
    p.noSlip, p.noBreak, p.noFrict, p.noAttraction = int(noSlip), int(noBreak), int(noFrict), int(noAttraction)
    p.trackEnergy, p.trsfRenorm, p.dist00, p.loneMask = int(trackEnergy), trsfRenorm, int(dist00), loneMask
    p.precision = precision
    return p


def params_from_meta(a, trackEnergy=False, precision=PREC_FP64):
    m = a.meta
    law = LAW_HERTZ_DMT if m.get("law") == "Hertz" else LAW_IDEALELPL
    return make_params(dt=m["dt"], damping=m.get("damping", 0.), gravity=m.get("gravity", (0, 0, 0)), law=law,
                       poisson=m.get("poisson", .2), en=m.get("en", float("nan")), trackEnergy=trackEnergy,
                       precision=precision)


def cell_from_meta(a):
    m = a.meta
    if not m.get("periodic"):
        return dict(periodic=False, hSize=np.eye(3), gradV=np.zeros((3, 3)), nextGradV=np.zeros((3, 3)), homoDeform=4)
    L = m["L"]
    sr = m.get("strainRate", 0.)
    return dict(periodic=True, hSize=np.diag([L, L, L]), gradV=np.zeros((3, 3)), nextGradV=sr * np.eye(3), homoDeform=4)


class HipRun:
    """One scene on one context: upload once, step, read back."""

    def __init__(self, arrays, params, cell, ctx=None):
        self.ctx = ctx or Context(0)
        self.a = arrays.normalized()
        self.params, self.cell0 = params, cell
        self.sid = self.ctx.scene_alloc(self.a.n_par, self.a.n_node, self.a.n_mat)
        self.ctx.upload(self.sid, self.a)
        self.ctx.set_cell(self.sid, cell["periodic"], cell["hSize"], cell["gradV"], cell["nextGradV"], cell["homoDeform"])
        self.ctx.set_params(self.sid, params)

    def step(self, n=1):
        self.ctx.step([self.sid], n)

    def nodes(self):
        return self.ctx.download_nodes(self.sid, self.a.n_node)

    def contacts(self):
        return self.ctx.contacts(self.sid)

    def energy(self):
        return self.ctx.energy(self.sid)

    def stats(self):
        return self.ctx.stats(self.sid)

    def cell(self):
        return self.ctx.cell(self.sid)

    def close(self):
        if self.sid is not None:
            self.ctx.scene_free(self.sid)
            self.sid = None
