"""Serving: continuous batching of DEM scenes on one device.

Requests (a scene, its engine attributes and a number of steps) arrive over time.  A
request is admitted when a slot is free and its HBM footprint (wdm_scene_footprint)
fits next to the resident scenes; admitted scenes stay device-resident and are stepped
together, one wdm_step call per chunk of steps for all of them (one launch per engine
for the whole batch).  Requests arriving meanwhile join at the next chunk boundary,
finished ones are read back and freed.  Each scene's trajectory does not depend on
what else is in the batch (per-scene block layout and reduction order), so a result
is the same alone and in the mix.

A failing request is confined to itself, as a failing Woo scene is to its run thread
(core/Scene.cpp:90-95): a request larger than the whole budget, or whose upload fails, is
rejected on its own (its slot freed); a scene stopped by a device error during a batched
step (NaN force, oversized particle) is evicted with its error while the others, which
took their steps, go on.  Such requests finish with `error` set and no result.

The scene of each request is built by the client (the payload); the server's clock
covers admission, upload, stepping and read-back.
"""
import math
import time

import numpy as np


class Request:
    def __init__(self, rid, arrays, params, cell, steps, arrival):
        self.rid, self.arrays, self.params, self.cell, self.steps, self.arrival = rid, arrays, params, cell, int(steps), float(arrival)
        self.sid = None
        self.done_steps = 0
        self.admitted = self.finished = None
        self.footprint = 0
        self.result = None
        self.error = None  # why the request ended without a result

    @property
    def particles(self):
        return self.arrays.n_par


def poisson_arrivals(n, rate, seed):
    """Arrival times of n requests with exponential gaps of mean 1/rate (seeded)."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1. / rate, n))


class Server:
    """Continuous-batching loop over one wdm context.

    ctx: woodem_amd._native.Context (or an object with the same scene_alloc/upload/
    set_cell/set_params/step/scene_error/download_nodes/scene_free/footprint methods);
    hbm_budget: bytes the resident scenes may occupy; max_slots: resident scenes at most;
    chunk: steps per batched call at most; every resident scene steps in every call, a
    call being shortened to the fewest steps a resident request has left, so every request
    runs exactly its steps and none waits for the others.
    contacts_per_particle: the contact density assumed for the footprint."""

    def __init__(self, ctx, hbm_budget, max_slots=1024, chunk=100, contacts_per_particle=8., clock=time.perf_counter,
                 sleep=time.sleep, log=None, log_every=5.):
        self.ctx, self.budget, self.max_slots, self.chunk = ctx, int(hbm_budget), int(max_slots), int(chunk)
        self.cpp = float(contacts_per_particle)
        self.clock, self.sleep = clock, sleep
        self.resident = []
        self.used = 0
        self.queue = []  # admitted in arrival order
        self.finished = []
        self.chunk_ms = []  # wall time per batched step, one entry per chunk
        self.particle_steps = 0
        self.max_resident = 0
        self.rejected = self.evicted = 0
        self.t0 = None
        self.log, self.log_every = log, float(log_every)

    # ---- admission ----
    def footprint(self, req):
        a = req.arrays
        return int(self.ctx.footprint(a.n_par, a.n_node, self.cpp))

    def _reject(self, req, why):
        req.error = why
        req.finished = self.clock() - self.t0
        self.finished.append(req)
        self.rejected += 1
        if self.log:
            self.log("request %r rejected: %s" % (req.rid, why))

    def _admit(self, now):
        while self.queue and self.queue[0].arrival <= now:
            req = self.queue[0]
            fp = self.footprint(req)
            if fp > self.budget:
                self.queue.pop(0)
                self._reject(req, "needs %d bytes of HBM, more than the whole budget of %d" % (fp, self.budget))
                continue
            if len(self.resident) >= self.max_slots or self.used + fp > self.budget:
                return  # head of line waits for memory or a slot (arrival order is kept)
            self.queue.pop(0)
            try:
                req.sid = self.ctx.scene_alloc(req.arrays.n_par, req.arrays.n_node, req.arrays.n_mat)
                self.ctx.upload(req.sid, req.arrays)
                c = req.cell
                self.ctx.set_cell(req.sid, c["periodic"], c["hSize"], c["gradV"], c["nextGradV"], c["homoDeform"])
                self.ctx.set_params(req.sid, req.params)
            except Exception as e:  # this request only: its slot is freed, the others carry on
                if req.sid is not None:
                    self.ctx.scene_free(req.sid)
                    req.sid = None
                self._reject(req, "admission failed: %s" % e)
                continue
            req.footprint = fp
            req.admitted = self.clock() - self.t0
            self.resident.append(req)
            self.used += fp
            self.max_resident = max(self.max_resident, len(self.resident))

    def _retire(self, req, error=None):
        if error is None:
            req.result = self.ctx.download_nodes(req.sid, req.arrays.n_node)
        else:
            req.error = error
            self.evicted += 1
            if self.log:
                self.log("request %r evicted after %d steps: %s" % (req.rid, req.done_steps, error))
        self.ctx.scene_free(req.sid)
        req.sid = None
        req.finished = self.clock() - self.t0
        self.used -= req.footprint
        self.resident.remove(req)
        self.finished.append(req)

    def _step(self, reqs, n):
        t = self.clock()
        failed = {}
        try:
            self.ctx.step([r.sid for r in reqs], n)
        except Exception as e:
            # the scenes without an error took their n steps; the stopped ones are evicted
            for r in reqs:
                code, msg = self.ctx.scene_error(r.sid)
                if code:
                    failed[r.rid] = "%s (%d)" % (msg, code)
            if not failed:
                raise e
        dt = self.clock() - t
        self.chunk_ms.append(dt / n * 1e3)
        for r in reqs:
            if r.rid in failed:
                self._retire(r, error=failed[r.rid])
                continue
            r.done_steps += n
            self.particle_steps += r.particles * n

    # ---- main loop ----
    def serve(self, requests):
        """Run every request to completion; arrival times are seconds after the start."""
        self.queue = sorted(requests, key=lambda r: r.arrival)
        self.t0 = self.clock()
        next_log = self.log_every
        while self.queue or self.resident:
            now = self.clock() - self.t0
            if self.log and now >= next_log:
                self.log("t=%.1f s: %d resident, %d finished, %d waiting, %.3g particle-steps" % (
                    now, len(self.resident), len(self.finished), len(self.queue), self.particle_steps))
                next_log = now + self.log_every
            self._admit(now)
            if not self.resident:
                wait = self.queue[0].arrival - now
                if wait > 0:
                    self.sleep(wait)
                continue
            # every resident scene steps; the call ends where the first of them is done
            n = min(self.chunk, min(r.steps - r.done_steps for r in self.resident))
            self._step(list(self.resident), n)
            for r in [r for r in self.resident if r.done_steps >= r.steps]:
                self._retire(r)
        return self.finished

    def report(self):
        wall = (self.clock() - self.t0) if self.t0 is not None else float("nan")
        ok = [r for r in self.finished if r.error is None]
        lat = np.array([r.finished - r.arrival for r in ok])
        wait = np.array([r.admitted - r.arrival for r in ok])
        cms = np.array(self.chunk_ms) if self.chunk_ms else np.array([math.nan])
        return dict(requests=len(ok), rejected=self.rejected, evicted=self.evicted, wall_s=wall, particle_steps=self.particle_steps,
                    particle_steps_per_s=self.particle_steps / wall if wall > 0 else math.nan,
                    latency_s={"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))} if len(lat) else None,
                    admission_wait_s={"p50": float(np.percentile(wait, 50)), "p99": float(np.percentile(wait, 99))} if len(wait) else None,
                    step_ms={"p50": float(np.percentile(cms, 50)), "p99": float(np.percentile(cms, 99))},
                    max_resident=self.max_resident)
