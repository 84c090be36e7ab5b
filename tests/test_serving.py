"""Serving loop (woodem_amd.serving): admission by HBM footprint and slots, continuous
batching, latency bookkeeping (CPU, with a stand-in context and a simulated clock), and
on the GPU: every request's result is the same alone and in the mix."""
import numpy as np
import pytest

from woodem_amd import runner, scenes, serving


class FakeCtx:
    """Stand-in for woodem_amd._native.Context: records residency, advances a clock."""

    def __init__(self, clock, ms_per_particle_step=1e-6):
        self.clock = clock
        self.k = ms_per_particle_step
        self.res = {}
        self.next = 0
        self.calls = []

    @staticmethod
    def footprint(n_par, n_node, cpp):
        return 1000 * n_par

    def scene_alloc(self, n_par, n_node, n_mat):
        self.next += 1
        self.res[self.next] = n_par
        return self.next

    def upload(self, sid, a):
        pass

    def set_cell(self, sid, *a):
        pass

    def set_params(self, sid, p):
        pass

    fail_at = {}  # sid -> number of the step call in which that scene stops with a device error
    failed = ()

    def step(self, sids, n):
        assert len(set(sids)) == len(sids) and all(s in self.res for s in sids)
        assert not any(s in self.failed for s in sids), "a stopped scene was stepped again"
        self.calls.append((tuple(sids), n))
        self.clock.t += n * self.k * sum(self.res[s] for s in sids) * 1e-3
        bad = [s for s in sids if self.fail_at.get(s) == len(self.calls)]
        if bad:
            self.failed = tuple(self.failed) + tuple(bad)
            raise RuntimeError("WDM_E_NAN (-4): scene(s) %s stopped by a device error" % bad)

    def scene_error(self, sid):
        return (-4, "NaN force in contact") if sid in self.failed else (0, "")

    def download_nodes(self, sid, n_node):
        return {"sid": sid}

    def scene_free(self, sid):
        del self.res[sid]


class Clock:
    def __init__(self):
        self.t = 0.

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += dt


class Arr:
    def __init__(self, n):
        self.n_par = self.n_node = n
        self.n_mat = 1


def reqs(sizes, arrivals, steps=1000):
    return [serving.Request(i, Arr(n), None, dict(periodic=False, hSize=None, gradV=None, nextGradV=None, homoDeform=4), steps, t)
            for i, (n, t) in enumerate(zip(sizes, arrivals))]


def test_admission_respects_budget_and_slots():
    clk = Clock()
    ctx = FakeCtx(clk)
    rs = reqs([1000, 2000, 3000, 500, 4000], [0., .001, .002, .003, .004])
    srv = serving.Server(ctx, hbm_budget=1000 * 5000, max_slots=2, chunk=100, clock=clk, sleep=clk.sleep)
    peak = {"n": 0, "mem": 0}
    orig = ctx.step

    def step(sids, n):
        peak["n"] = max(peak["n"], len(ctx.res))
        peak["mem"] = max(peak["mem"], sum(ctx.res.values()) * 1000)
        orig(sids, n)

    ctx.step = step
    done = srv.serve(rs)
    assert len(done) == 5 and all(r.done_steps == 1000 for r in done)
    assert peak["n"] <= 2 and peak["mem"] <= 1000 * 5000
    for r in done:
        assert r.admitted >= r.arrival - 1e-12 and r.finished > r.admitted
    # first-come first-served admission
    adm = sorted(done, key=lambda r: r.admitted)
    assert [r.rid for r in adm] == sorted(r.rid for r in adm)
    rep = srv.report()
    assert rep["requests"] == 5 and rep["particle_steps"] == 1000 * sum([1000, 2000, 3000, 500, 4000])


def test_continuous_batching_joins_at_chunk_boundaries():
    clk = Clock()
    ctx = FakeCtx(clk, ms_per_particle_step=1e-3)
    # the second request arrives while the first runs: it must join the running batch
    rs = reqs([1000, 1000], [0., .05], steps=2000)
    srv = serving.Server(ctx, hbm_budget=10 ** 9, chunk=100, clock=clk, sleep=clk.sleep)
    srv.serve(rs)
    both = [c for c in ctx.calls if len(c[0]) == 2]
    assert both, "requests never shared a batched call"
    assert all(n == 100 for _, n in ctx.calls)


def test_uneven_step_counts_finish_exactly():
    clk = Clock()
    ctx = FakeCtx(clk)
    rs = reqs([100, 200, 300], [0., 0., 0.], steps=250)
    srv = serving.Server(ctx, hbm_budget=10 ** 9, chunk=100, clock=clk, sleep=clk.sleep)
    done = srv.serve(rs)
    assert all(r.done_steps == 250 for r in done)


def test_oversized_request_rejected_alone():
    """a request larger than the whole budget ends with an error; the others are served"""
    clk = Clock()
    ctx = FakeCtx(clk)
    srv = serving.Server(ctx, hbm_budget=10000 * 1000, chunk=10, clock=clk, sleep=clk.sleep)
    done = srv.serve(reqs([10, 20000, 30], [0., 0., .001], steps=50))
    by = {r.rid: r for r in done}
    assert by[1].error and "budget" in by[1].error and by[1].result is None and by[1].sid is None
    assert by[0].error is None and by[2].error is None and by[0].done_steps == by[2].done_steps == 50
    assert not ctx.res  # every slot freed
    rep = srv.report()
    assert rep["requests"] == 2 and rep["rejected"] == 1 and rep["evicted"] == 0


def test_failed_upload_frees_its_slot():
    clk = Clock()
    ctx = FakeCtx(clk)
    orig = ctx.upload

    def upload(sid, a):
        if a.n_par == 77:
            raise RuntimeError("WDM_E_ARG (-1): node arrays")
        orig(sid, a)

    ctx.upload = upload
    srv = serving.Server(ctx, hbm_budget=10 ** 9, chunk=10, clock=clk, sleep=clk.sleep)
    done = srv.serve(reqs([10, 77, 30], [0., 0., 0.], steps=40))
    by = {r.rid: r for r in done}
    assert "admission failed" in by[1].error and by[0].done_steps == by[2].done_steps == 40
    assert not ctx.res


def test_scene_stopped_by_device_error_is_evicted_alone():
    """one scene of a batched call stops (NaN force): it is evicted with its error after the
    steps it took; the others, which took that call's steps too, run to completion"""
    clk = Clock()
    ctx = FakeCtx(clk)
    ctx.fail_at = {2: 3}  # scene id 2 (the second request) stops in the third call
    srv = serving.Server(ctx, hbm_budget=10 ** 9, chunk=10, clock=clk, sleep=clk.sleep)
    done = srv.serve(reqs([10, 20, 30], [0., 0., 0.], steps=100))
    by = {r.rid: r for r in done}
    assert by[1].error.startswith("NaN force in contact") and by[1].result is None and by[1].done_steps == 20
    assert by[0].done_steps == by[2].done_steps == 100 and by[0].result and by[2].result
    assert all(2 not in c[0] for c in ctx.calls[3:])
    assert not ctx.res
    rep = srv.report()
    assert rep["requests"] == 2 and rep["evicted"] == 1
    assert rep["particle_steps"] == 100 * 10 + 100 * 30 + 20 * 20


def drum_request(rid, seed, n, arrival, steps):
    d = scenes.drum(seed=seed, n=n, gap=1e-4)
    ns = d.meta["n_spheres"]
    d.vel[:ns, 2] = -.5
    return serving.Request(rid, d, runner.params_from_meta(d), runner.cell_from_meta(d), steps, arrival)


@pytest.mark.gpu
def test_served_results_equal_alone(hip_ctx):
    """Drums arriving at different times share batched steps; each final state equals
    the same drum stepped alone (bit for bit)."""
    clk = Clock()
    specs = [(0, 21, 300, 0.), (1, 22, 150, 0.), (2, 23, 500, .002), (3, 24, 220, .004)]
    steps = 600
    mk = lambda: [drum_request(i, s, n, t, steps) for i, s, n, t in specs]
    orig_step = hip_ctx.step

    def timed_step(sids, n):  # deterministic simulated time: 10 us per step
        orig_step(sids, n)
        clk.t += n * 1e-5

    hip_ctx.step = timed_step
    try:
        srv = serving.Server(hip_ctx, hbm_budget=8 << 30, chunk=100, clock=clk, sleep=clk.sleep)
        done = srv.serve(mk())
    finally:
        hip_ctx.step = orig_step
    assert srv.max_resident >= 3  # the requests overlapped on the device
    for r, alone in zip(sorted(done, key=lambda r: r.rid), mk()):
        g = runner.HipRun(alone.arrays, alone.params, alone.cell, ctx=hip_ctx)
        g.step(steps)
        n = g.nodes()
        g.close()
        for f in ("pos", "vel", "angVel", "ori", "force", "torque"):
            assert np.array_equal(r.result[f], n[f]), (r.rid, f)


@pytest.mark.gpu
def test_nan_drum_is_evicted_and_the_others_equal_alone(hip_ctx):
    """A drum whose material has a NaN Young's modulus gets NaN contact forces at its first
    contact (IdealElPl.cpp:76-80 raises there): it is evicted with WDM_E_NAN, and the drums
    stepped in the same calls end bit for bit as when stepped alone."""
    clk = Clock()
    steps = 400
    specs = [(0, 31, 300, 0.), (1, 32, 200, 0.), (2, 33, 250, 0.)]

    def mk():
        rs = [drum_request(i, s, n, t, steps) for i, s, n, t in specs]
        bad = rs[1].arrays
        bad.young[:] = np.nan
        bad.pos[1] = bad.pos[0] + np.array([0., 0., .5 * (bad.radius[0] + bad.radius[1])])  # overlapping at the start
        return rs

    orig_step = hip_ctx.step

    def timed_step(sids, n):
        try:
            orig_step(sids, n)
        finally:
            clk.t += n * 1e-5

    hip_ctx.step = timed_step
    try:
        srv = serving.Server(hip_ctx, hbm_budget=8 << 30, chunk=50, clock=clk, sleep=clk.sleep)
        done = srv.serve(mk())
    finally:
        hip_ctx.step = orig_step
    by = {r.rid: r for r in done}
    assert by[1].error and "NaN" in by[1].error and by[1].result is None and by[1].done_steps < steps
    assert srv.report()["evicted"] == 1
    for rid in (0, 2):
        alone = mk()[rid]
        g = runner.HipRun(alone.arrays, alone.params, alone.cell, ctx=hip_ctx)
        g.step(steps)
        n = g.nodes()
        g.close()
        for f in ("pos", "vel", "angVel", "ori", "force", "torque"):
            assert np.array_equal(by[rid].result[f], n[f]), (rid, f)
