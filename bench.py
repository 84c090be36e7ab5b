#!/usr/bin/env python3
"""Benchmark: particle-steps/s of the Woo DEM timestep on one MI355X.

Workload (default): 1,000,000 spheres in a periodic cell at 0.60 solid fraction,
lognormal radii (median 1 mm, sigma 0.3, clipped to [0.5, 3] mm), Hertz contact
(E=7e10, nu=.3, en=.7, tanPhi=.4), Gaussian velocities (std 0.05 m/s), no
gravity, isotropic cell strain rate 1e-3/s, FP64.  Overlaps of the uniform
placement are relaxed with 500 device sweeps before timing.  A step is one
Scene::doOneStep: Leapfrog, InsertionSortCollider (Verlet-triggered), ContactLoop.

--workload batch: 512 independent deposition boxes (woodem_amd.scenes.deposition_box,
seeds 1000..1511: N_s = round(exp(U(ln 500, ln 2e4))), r ~ U[2, 6] mm, tanPhi ~
U[.1, .8], damping ~ U[.05, .4]), FP64 state with FP32 contact arithmetic, stepped
together by one wdm_step call per step batch; the same JSON line carries the
sequential rate (each scene stepped alone, on a sample of the scenes) and the ratio.

--workload serve: a Poisson stream (20 requests/s, seeded) of rotating drums
(woodem_amd.scenes.drum: 1,024 facets + 2 end walls clumped to a spinning centre node,
1e3..2e5 spheres log-uniform capped at the drum's lattice capacity, r ~ U[1, 4] mm,
IdealElPl, FP64), 10,000 steps per request, served by woodem_amd.serving.Server
(admission by HBM footprint, continuous batching); value = aggregate particle-steps/s
over the serving wall time, with per-request latency p50/p99.

One JSON line on stdout.  --gpus N with N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# algorithmic HBM bytes per unit of the contact kernel (FP64 / mixed): real contact,
# potential (non-real) contact
BYTES_REAL = {"fp64": 496, "mixed": 324}
BYTES_POT = {"fp64": 64, "mixed": 60}
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


CPU_LABEL = "Woo-semantics CPU restatement (oracle/, C++17 + OpenMP)"
# a one-GPU box grants a job 16 host CPUs, whatever its affinity mask or nproc show
CPU_SHARE = 16


def cpu_setup():
    """Threads of the host baseline and the OpenMP placement, set before the OpenMP runtime
    starts (the oracle library is loaded after this): every CPU of the affinity mask up to the
    box's share, threads bound close to each other on cores."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(aff, CPU_SHARE))
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(threads=threads, affinity_cpus=aff, cpu_model=model, omp_proc_bind=os.environ["OMP_PROC_BIND"],
                omp_places=os.environ["OMP_PLACES"])


def oracle_rate(OL, arrays, params, cell, threads, seconds, warm=10, chunk=5):
    """particle-steps/s of the oracle stepping `arrays` on `threads` threads: `warm` untimed
    steps (the initial sort), then chunks of steps until `seconds` have passed"""
    o = OL.OracleRun(arrays, params, cell, nThreads=threads)
    o.step(warm)
    t0 = time.perf_counter()
    k = 0
    while True:
        o.step(chunk)
        k += chunk
        el = time.perf_counter() - t0
        if el > seconds:
            break
    o.close()
    return arrays.n_par * k / el, k, el


def cpu_baseline(ctx, n_sample, seconds, cpu):
    """The oracle on a bounded sample of the timed workload (same generator, PSD, packing
    fraction, relaxation and numbering, fewer spheres), on every thread of the box's share and
    on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as OL
    from woodem_amd import runner, scenes
    b = scenes.relax_and_order(scenes.periodic_hertz(n=n_sample), ctx, iters=500)
    p, cell = runner.params_from_meta(b), runner.cell_from_meta(b)
    v, k, el = oracle_rate(OL, b, p, cell, cpu["threads"], seconds)
    v1, k1, el1 = oracle_rate(OL, b, p, cell, 1, seconds / 3)
    return dict(value=v, unit="particle-steps/s", cores=cpu["threads"], kind="port", label=CPU_LABEL,
                value_1_thread=v1, cpu_model=cpu["cpu_model"], affinity_cpus=cpu["affinity_cpus"],
                omp={"OMP_PROC_BIND": cpu["omp_proc_bind"], "OMP_PLACES": cpu["omp_places"]},
                sample="periodic Hertz, %d spheres (the timed workload's generator, PSD, phi=0.60, relaxed), %d steps in "
                       "%.1f s after 10 warm-up steps (1 thread: %d steps in %.1f s)" % (n_sample, k, el, k1, el1))


def cloud_args(num):
    """SpherePack.makeCloud workload: num spheres, rRelFuzz .3, porosity .7 (mean radius from
    the porosity), in a cube of 0.27 m per 100,000 spheres"""
    L = 0.27 * (num / 1e5) ** (1. / 3.)
    return np.zeros(3), np.full(3, L), dict(rRelFuzz=.3, num=num, porosity=.7)


def cpu_make_cloud(n_sample):
    """the oracle's makeCloud (all placed spheres tested per try, as SpherePack.cpp:222-231)"""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as OL
    mn, mx, kw = cloud_args(n_sample)
    t = time.perf_counter()
    _, r, _, _ = OL.make_cloud(mn, mx, state=7, cap=n_sample, **kw)
    el = time.perf_counter() - t
    return dict(value=len(r) / el, unit="spheres/s", cores=1, kind="port",
                sample="makeCloud of %d spheres (same density; the cost per sphere grows with the count), %.2f s" % (len(r), el))


def gpu_make_cloud(ctx, num):
    mn, mx, kw = cloud_args(num)
    st = np.array([7], np.uint32)
    ctx.make_cloud(mn, mx, -1., kw["rRelFuzz"], 2000, False, kw["porosity"], [], [], False, st, num)  # warm-up
    st = np.array([7], np.uint32)
    ctx.sync()
    t = time.perf_counter()
    _, r, _ = ctx.make_cloud(mn, mx, -1., kw["rRelFuzz"], num, False, kw["porosity"], [], [], False, st, num)
    el = time.perf_counter() - t
    return dict(value=len(r) / el, unit="spheres/s", seconds=el, spheres=len(r),
                workload="SpherePack.makeCloud, num=%d, rRelFuzz .3, porosity .7, %.3f m cube, generator state 7" % (num, mx[0]))


def scene_io(ctx, run):
    """Checkpoint (woodem_amd.snapshot npz) save and restore, and VTK export, of the timed scene;
    host files under a temporary folder that is removed afterwards."""
    import shutil
    import tempfile
    from woodem_amd import snapshot
    d = tempfile.mkdtemp(prefix="wdm_bench_")
    try:
        ctx.sync()
        t = time.perf_counter()
        f = snapshot.save(run, os.path.join(d, "scene.npz"))
        ts = time.perf_counter() - t
        size = os.path.getsize(f)
        t = time.perf_counter()
        r2 = snapshot.load(f, ctx=ctx)
        ctx.sync()
        tl = time.perf_counter() - t
        r2.close()
        t = time.perf_counter()
        files = snapshot.vtk_export(run, os.path.join(d, "vtk", "scene."))
        tv = time.perf_counter() - t
        vsize = sum(os.path.getsize(x) for x in files.values())
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return {"checkpoint": {"bytes": size, "save_s": ts, "load_s": tl, "save_GBps": size / ts / 1e9, "load_GBps": size / tl / 1e9,
                           "what": "nodes, particles, materials, contacts with their state, energy, cell, engine state (npz)"},
            "vtk": {"bytes": vsize, "s": tv, "files": sorted(files), "format": "VTK XML, inline base64 binary"}}


def insertion(ctx, n_spare=20000):
    """Particles inserted in place (DemField.par.add during a run, ParticleContainer ids): a
    200,000-sphere box (100 x 100 x 20 lattice, r 5 mm, gravity, walls) with spare slots gets
    n_spare spheres in one wdm_insert_spheres call (two lattice layers above the packing); the
    step after it runs a full collider pass (the new particles need bounds).  Also the
    BoxOutlet selection (wdm_select_box) over the box's lower half."""
    from woodem_amd import runner, scenes, soa
    a = scenes.box_lattice(nx=100, ny=100, nz=20, box=((0., 0., 0.), (1.2, 1.2, 1.0)))
    a = soa.with_spare(a, n_spare)
    p = runner.params_from_meta(a)
    run = runner.HipRun(a, p, runner.cell_from_meta(a), ctx=ctx)
    run.step(20)
    ctx.sync()
    t = time.perf_counter()
    run.step(1)
    ctx.sync()
    t_step = time.perf_counter() - t
    r, rho = 5e-3, 3000.
    m = 4. / 3. * math.pi * r ** 3 * rho
    inertia = .4 * m * r * r
    g = np.arange(100) * 11e-3 + .055
    xx, yy, zz = np.meshgrid(g, g, [.6, .611], indexing="ij")
    pos = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], 1)[:n_spare]
    t = time.perf_counter()
    ids = ctx.insert_spheres(run.sid, pos, r, m, inertia, 0, soa.DEFAULT_MOVABLE_MASK)
    ctx.sync()
    t_ins = time.perf_counter() - t
    full0 = run.stats()["nFullRuns"]
    t = time.perf_counter()
    run.step(1)
    ctx.sync()
    t_after = time.perf_counter() - t
    full1 = run.stats()["nFullRuns"]
    t = time.perf_counter()
    sel = ctx.select_box(run.sid, (0., 0., 0.), (1.2, 1.2, .15), inside=True)
    t_sel = time.perf_counter() - t
    run.close()
    return {"scene": "box_lattice 100x100x20 (200,000 spheres, 6 walls) + %d spare slots" % n_spare, "inserted": int(len(ids)),
            "insert_s": t_ins, "inserted_per_s": len(ids) / t_ins, "step_before_s": t_step, "step_after_s": t_after,
            "full_collider_runs_in_step_after": full1 - full0, "select_box_s": t_sel, "selected": int(len(sel))}


def measured_traffic(args, n_real, n_active):
    """HBM bytes per k_contacts launch (read + write) from the rocprofv3 FETCH_SIZE/WRITE_SIZE
    runs in profiles/, when they were taken at this run's roofline steps (the same real and
    processed contact counts: the steps are deterministic); otherwise None, with the reason."""
    path = os.path.join(ROOT, "profiles", "traffic_periodic_hertz_1M_fp64.json")
    if args.n != 1_000_000 or args.precision != "fp64" or not os.path.exists(path):
        return None, "no counter profile for this workload"
    with open(path) as f:
        t = json.load(f)
    if t.get("real_contacts") != n_real or t.get("processed_contacts") != n_active:
        return None, "the counter profile was taken at other steps (%s real, %s processed contacts)" % (
            t.get("real_contacts"), t.get("processed_contacts"))
    return t["hbm_read_bytes_per_launch"] + t["hbm_write_bytes_per_launch"], "profiles/" + t.get("source", "")


COPY_PROBE = r"""
import json, sys, torch
nbytes, reps = int(sys.argv[1]), int(sys.argv[2])
a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
b = torch.empty_like(a)
b.copy_(a)
torch.cuda.synchronize()
e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
e0.record()
for _ in range(reps):
    b.copy_(a)
e1.record()
torch.cuda.synchronize()
print(json.dumps({"ms_per_copy": e0.elapsed_time(e1) / reps}))
"""


def hbm_copy_rate(nbytes=1 << 30, reps=10):
    """A plain device-to-device copy of nbytes (torch, in a child process of its own: torch's HIP
    runtime does not start in a process that has loaded the library's), timed with events:
    the HBM rate a streaming kernel reaches on this device, beside the nominal peak."""
    import subprocess
    r = subprocess.run([sys.executable, "-c", COPY_PROBE, str(nbytes), str(reps)], capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        raise RuntimeError("copy probe failed: " + r.stderr[-300:])
    ms = json.loads(r.stdout.strip().splitlines()[-1])["ms_per_copy"]
    gbps = 2 * nbytes / (ms * 1e-3) / 1e9
    return {"bytes_moved_per_copy": 2 * nbytes, "ms_per_copy": ms, "GBps": gbps, "frac_of_peak": gbps / HBM_PEAK_GBS,
            "how": "torch copy_ of 1 GiB float64, 10 repetitions between two events, in a child process"}


def timed_kernels(ctx, sids, steps):
    """HIP-event kernel times over `steps` steps launched one at a time (window 1), so that
    every launch of a step is counted once."""
    t = ctx.tuning()
    ctx.set_tuning(window=1)
    ctx.enable_timing(True)
    ctx.step(sids, steps)
    kt = ctx.kernel_times()
    ctx.enable_timing(False)
    ctx.set_tuning(window=t["window"])
    return kt


def launch_config(ctx, tune):
    t = ctx.tuning()
    out = {"window_steps": t["window"], "graphs": bool(t["graphs"]), "cell_factor": t["cell_factor"], "block": 256}
    if tune:
        out["autotuned"] = True
        out["autotune_steps"] = tune["steps"]
        out["autotune_trials"] = [[x["window"], x["cell_factor"], round(x["ms_per_step"], 4)] for x in tune["trials"]]
    else:
        out["autotuned"] = False
    return out


def bench_batch(args, ctx):
    """Many small scenes stepped as one batch vs the same scenes stepped one by one."""
    from woodem_amd import runner, scenes
    from woodem_amd._native import PREC_MIXED
    t0 = time.time()
    arrs = [scenes.relax_box(scenes.deposition_box(seed=1000 + s), ctx, iters=args.relax_box) for s in range(args.scenes)]
    runs = [runner.HipRun(a, runner.params_from_meta(a, precision=PREC_MIXED), runner.cell_from_meta(a), ctx=ctx) for a in arrs]
    sids = [r.sid for r in runs]
    n_tot = sum(a.n_par for a in arrs)
    log("batch: %d scenes, %d particles in total (%d..%d per scene), set up in %.1f s" % (
        len(arrs), n_tot, min(a.n_par for a in arrs), max(a.n_par for a in arrs), time.time() - t0))
    tune = None
    if args.autotune:
        # tuned on a copy of the batch (same scenes, own state), so that the timed steps
        # below start from the same state as an untuned run
        from woodem_amd.autotune import autotune
        tt = time.perf_counter()
        twins = [runner.HipRun(a, runner.params_from_meta(a, precision=PREC_MIXED), runner.cell_from_meta(a), ctx=ctx) for a in arrs]
        tsids = [r.sid for r in twins]
        ctx.step(tsids, args.warmup)
        tune = autotune(ctx, tsids, steps=args.tune_steps, warm=10, windows=(4, 8, 16, 32, 64, 128), log=log)
        for r in twins:
            r.close()
        log("autotune: window %d, cell factor %.2f (%d steps on a copy of the batch, %.1f s)" % (
            tune["window"], tune["cell_factor"], tune["steps"], time.perf_counter() - tt))
    ctx.step(sids, args.warmup)
    ctx.sync()
    fr0 = sum(ctx.stats(s)["nFullRuns"] for s in sids)
    t1 = time.perf_counter()
    ctx.step(sids, args.steps)
    ctx.sync()
    el = time.perf_counter() - t1
    fr1 = sum(ctx.stats(s)["nFullRuns"] for s in sids)
    value = n_tot * args.steps / el
    log("batched: %d steps in %.3f s: %.4g particle-steps/s, %d scene full runs" % (args.steps, el, value, fr1 - fr0))
    # per-step latency of the whole batch
    lat = []
    for _ in range(args.latency_steps):
        tt = time.perf_counter()
        ctx.step(sids, 1)
        lat.append(time.perf_counter() - tt)
    lat = np.array(lat) * 1e3
    # kernel time per step (all launches of a kind in one step summed), one step at a time
    kt = timed_kernels(ctx, sids, args.timing_steps)
    n_con = n_real = n_active = 0
    for s in sids:
        c = ctx.contacts(s)
        n_con += len(c["idA"])
        n_real += int(c["isReal"].sum())
        n_active += ctx.contact_work(s)[0]
    ms_step = kt.get("contacts", (float("nan"), 0))[0] / args.timing_steps
    alg_bytes = BYTES_REAL["mixed"] * n_real + BYTES_POT["mixed"] * max(n_active - n_real, 0)
    achieved = alg_bytes / (ms_step * 1e-3) / 1e9
    kernels = {k: {"ms_per_step": v[0] / args.timing_steps, "launches": v[1]} for k, v in kt.items()}
    log("kernels: %s" % json.dumps(kernels))
    # sequential: a sample of the scenes, each stepped alone (same code path, batch of one)
    pick = list(range(0, len(runs), max(1, len(runs) // args.seq_scenes)))[:args.seq_scenes]
    seq_ps = 0
    seq_t = 0.
    for i in pick:
        r = runner.HipRun(arrs[i], runner.params_from_meta(arrs[i], precision=PREC_MIXED), runner.cell_from_meta(arrs[i]), ctx=ctx)
        r.step(args.warmup)
        ctx.sync()
        tt = time.perf_counter()
        r.step(args.seq_steps)
        ctx.sync()
        seq_t += time.perf_counter() - tt
        seq_ps += arrs[i].n_par * args.seq_steps
        r.close()
    seq = seq_ps / seq_t
    log("sequential sample: %d scenes x %d steps: %.4g particle-steps/s; batched/sequential = %.2f" % (
        len(pick), args.seq_steps, seq, value / seq))
    for r in runs:
        r.close()
    return {
        "metric": "particle-steps/s on 1 MI355X (batched scenes vs sequential); p50/p99 step latency",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64-state/f32-contact",
        "data": "synthetic (seeded generator, seeds 1000..%d)" % (999 + len(arrs)),
        "config": {"workload": "%d deposition boxes, %d particles in total, FP32 contact arithmetic, one batch" % (len(arrs), n_tot),
                   "scenes": len(arrs), "n_particles": n_tot, "precision": "mixed",
                   "contacts_potential": n_con, "contacts_real": n_real, "scene_full_runs_in_timed_region": fr1 - fr0,
                   "launch": launch_config(ctx, tune)},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)), "samples": len(lat)},
        "roofline": {"bound": "hbm", "kernel": "k_contacts", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_step": alg_bytes, "ms_per_step": ms_step},
        "kernels": kernels,
        "sequential": {"value": seq, "unit": "particle-steps/s", "scenes": len(pick), "steps": args.seq_steps,
                       "batched_over_sequential": value / seq},
        "cpu_baseline": None,
    }


def bench_serve(args, ctx):
    """Continuous batching of rotating-drum requests arriving as a Poisson stream."""
    from woodem_amd import runner, scenes, serving
    t0 = time.time()
    arr = serving.poisson_arrivals(args.requests, args.rate, seed=args.seed)
    reqs = []
    for i, t in enumerate(arr):
        d = scenes.drum(seed=args.seed * 100000 + i)
        reqs.append(serving.Request(i, d, runner.params_from_meta(d), runner.cell_from_meta(d), args.req_steps, float(t)))
    n_req = np.array([r.arrays.meta["n_requested"] for r in reqs])
    n_sph = np.array([r.arrays.meta["n_spheres"] for r in reqs])
    log("serve: %d drum requests over %.2f s, %d..%d spheres (mean %.0f; %d capped at the drum capacity), built in %.1f s" % (
        len(reqs), arr[-1], n_sph.min(), n_sph.max(), n_sph.mean(), int((n_req > n_sph).sum()), time.time() - t0))
    # warm-up: one small request through the same path (code objects, allocator)
    w = scenes.drum(seed=1, n=2000)
    serving.Server(ctx, args.hbm_budget, chunk=args.chunk).serve(
        [serving.Request(-1, w, runner.params_from_meta(w), runner.cell_from_meta(w), 200, 0.)])
    srv = serving.Server(ctx, args.hbm_budget, chunk=args.chunk, log=log)
    srv.serve(reqs)
    rep = srv.report()
    log("served: %s" % json.dumps(rep))
    # sequential: a sample of the requests, each stepped alone from its own state (no batching)
    pick = list(range(0, len(reqs), max(1, len(reqs) // args.seq_requests)))[:args.seq_requests]
    sps = st = 0.
    for i in pick:
        d = reqs[i].arrays
        r = runner.HipRun(d, reqs[i].params, reqs[i].cell, ctx=ctx)
        r.step(args.chunk)
        ctx.sync()
        tt = time.perf_counter()
        r.step(args.seq_steps)
        ctx.sync()
        st += time.perf_counter() - tt
        sps += d.n_par * args.seq_steps
        r.close()
    seq = {"value": sps / st, "unit": "particle-steps/s", "requests": len(pick), "steps": args.seq_steps,
           "served_over_sequential": rep["particle_steps_per_s"] / (sps / st)}
    log("serve sequential sample: %s" % json.dumps(seq))
    return {
        "metric": "particle-steps/s on 1 MI355X (served drums, continuous batching); p50/p99 latency",
        "value": rep["particle_steps_per_s"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.req_steps,
        "warmup": 200,
        "ms_per_step": rep["step_ms"]["p50"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Poisson arrivals and drum generator, seed %d)" % args.seed,
        "config": {"workload": "%d rotating-drum requests, Poisson %.0f/s, %d steps each, served with continuous batching" % (
            len(reqs), args.rate, args.req_steps), "requests": len(reqs), "rate_per_s": args.rate, "chunk_steps": args.chunk,
                   "spheres_mean": float(n_sph.mean()), "hbm_budget": args.hbm_budget, "launch": launch_config(ctx, None),
                   "rejected": rep["rejected"], "evicted": rep["evicted"]},
        "sequential": seq,
        "latency_s": rep["latency_s"],
        "admission_wait_s": rep["admission_wait_s"],
        "step_latency_ms": rep["step_ms"],
        "max_resident": rep["max_resident"],
        "wall_s": rep["wall_s"],
        "roofline": None,
        "cpu_baseline": None,
    }


def bench_cfg1(ctx, cpu, steps=2000, cpu_on=True):
    """Config 1 (10,000 spheres r 5 mm on a jittered 25 x 20 x 20 lattice in a box of 6 walls,
    IdealElPl, gravity, damping .2, dt 0.3 of the p-wave step, 2,000 steps) on the GPU in FP64
    and in mixed precision (config 3's speed is mixed over FP64 on this scene), and the oracle's
    rate on every thread of the box's share and on one thread, over the same 2,000 steps."""
    from woodem_amd import runner, scenes
    from woodem_amd._native import PREC_FP64, PREC_MIXED
    a = scenes.box_lattice()
    cell = runner.cell_from_meta(a)
    out = {"workload": "cfg1: box_lattice 25x20x20, %d particles (10,000 spheres + 6 walls), IdealElPl, %d steps" % (a.n_par, steps),
           "n_particles": a.n_par, "steps": steps}
    for name, prec in (("fp64", PREC_FP64), ("mixed", PREC_MIXED)):
        r = runner.HipRun(a, runner.params_from_meta(a, precision=prec), cell, ctx=ctx)
        r.step(1)  # first step: uploads, first collider run
        ctx.sync()
        t = time.perf_counter()
        r.step(steps - 1)
        ctx.sync()
        el = time.perf_counter() - t
        out[name] = {"value": a.n_par * (steps - 1) / el, "ms_per_step": el / (steps - 1) * 1e3,
                     "full_collider_runs": r.stats()["nFullRuns"]}
        r.close()
    out["cfg3_mixed_over_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    if cpu_on:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as OL
        p = runner.params_from_meta(a)
        res = {}
        for nt in (cpu["threads"], 1):
            o = OL.OracleRun(a, p, cell, nThreads=nt)
            o.step(1)
            t = time.perf_counter()
            o.step(steps - 1)
            el = time.perf_counter() - t
            o.close()
            res[nt] = a.n_par * (steps - 1) / el
        out["cpu_baseline"] = {"value": res[cpu["threads"]], "value_1_thread": res[1], "cores": cpu["threads"],
                               "unit": "particle-steps/s", "kind": "port", "label": CPU_LABEL, "cpu_model": cpu["cpu_model"]}
        out["gpu_fp64_over_cpu"] = out["fp64"]["value"] / res[cpu["threads"]]
    return out


def bench_periodic(args, ctx):
    """The default line: the 1M-sphere periodic Hertz scene in FP64 (or mixed), timed over
    exactly --steps steps after --warmup steps, with its roofline, latency, a sustained rate over
    the workload's 1,000 steps (collider full runs included), and the other configurations'
    legs."""
    from woodem_amd import runner, scenes
    from woodem_amd._native import PREC_FP64, PREC_MIXED
    cpu = cpu_setup()
    t0 = time.time()
    a = scenes.periodic_hertz(n=args.n)
    b = scenes.relax_and_order(a, ctx, iters=args.relax)
    log("scene: %d spheres, L=%.4f m, dt=%.3e s, relaxed in %.1f s" % (b.n_par, b.meta["L"], b.meta["dt"], time.time() - t0))
    prec = PREC_MIXED if args.precision == "mixed" else PREC_FP64
    p = runner.params_from_meta(b, precision=prec)
    run = runner.HipRun(b, p, runner.cell_from_meta(b), ctx=ctx)
    run.step(args.warmup)
    ctx.sync()
    log("warm-up done: %s" % run.stats())

    st0 = run.stats()
    t1 = time.perf_counter()
    run.step(args.steps)
    ctx.sync()
    el = time.perf_counter() - t1
    st1 = run.stats()
    value = b.n_par * args.steps / el
    log("timed %d steps in %.3f s: %.4g particle-steps/s, full collider runs %d" % (
        args.steps, el, value, st1["nFullRuns"] - st0["nFullRuns"]))

    # per-step latency (each call synchronises)
    lat = []
    for _ in range(args.latency_steps):
        tt = time.perf_counter()
        run.step(1)
        lat.append(time.perf_counter() - tt)
    lat = np.array(lat if lat else [np.nan]) * 1e3

    # kernel timing with HIP events on the library's stream, one step at a time
    kt = timed_kernels(ctx, [run.sid], args.timing_steps)
    c = run.contacts()
    n_con = len(c["idA"])
    n_real = int(c["isReal"].sum())
    n_active, _ = ctx.contact_work(run.sid)  # contacts the kernel reads (dormant ones it never touches)
    ms, launches = kt.get("contacts", (float("nan"), 0))
    avg_ms = ms / max(launches, 1)
    alg_bytes = BYTES_REAL[args.precision] * n_real + BYTES_POT[args.precision] * max(n_active - n_real, 0)
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic(args, n_real, n_active)
    kernels = {k: {"avg_ms": v[0] / max(v[1], 1), "launches": v[1]} for k, v in kt.items()}
    log("kernels: %s" % json.dumps(kernels))

    # sustained: the workload's 1,000 steps in windows, collider full runs included
    sus = None
    if args.sustained_steps > 0:
        f0 = run.stats()["nFullRuns"]
        ctx.sync()
        t = time.perf_counter()
        run.step(args.sustained_steps)
        ctx.sync()
        els = time.perf_counter() - t
        sus = {"value": b.n_par * args.sustained_steps / els, "steps": args.sustained_steps,
               "ms_per_step": els / args.sustained_steps * 1e3, "full_collider_runs": run.stats()["nFullRuns"] - f0}
        log("sustained: %s" % json.dumps(sus))

    # DynDt (utils.defaultEngines adds it with stepPeriod=100): its kernel on the same contacts
    ctx.enable_timing(True)
    for _ in range(10):
        crit = ctx.dyndt_crit(run.sid)
    kd = ctx.kernel_times().get("dyndt", (float("nan"), 0))
    ctx.enable_timing(False)
    kernels["dyndt"] = {"avg_ms": kd[0] / max(kd[1], 1), "launches": kd[1], "crit_dt": crit}
    gen = gpu_make_cloud(ctx, args.cloud_n) if args.cloud_n > 0 else None
    if gen:
        log("scene generation: %s" % json.dumps(gen))
    io = None
    if not args.no_io:
        io = scene_io(ctx, run)
        io["insertion"] = insertion(ctx)
        log("scene I/O and insertion: %s" % json.dumps(io))
    log("contacts: %d potential, %d real, %d processed by the ContactLoop" % (n_con, n_real, n_active))
    run.close()

    mixed = None
    if not args.no_mixed and args.precision == "fp64" and sus is not None:
        # the mixed-precision scene over the same steps as each FP64 figure: the timed steps
        # of the headline, the kernel-by-kernel steps of the roofline and the sustained steps
        pm = runner.params_from_meta(b, precision=PREC_MIXED)
        rm = runner.HipRun(b, pm, runner.cell_from_meta(b), ctx=ctx)
        rm.step(args.warmup)
        ctx.sync()
        tm = time.perf_counter()
        rm.step(args.steps)
        ctx.sync()
        elt = time.perf_counter() - tm
        rm.step(args.latency_steps)
        ktm = timed_kernels(ctx, [rm.sid], args.timing_steps)
        cm = rm.contacts()
        m_real, m_proc = int(cm["isReal"].sum()), rm.ctx.contact_work(rm.sid)[0]
        ctx.sync()
        f0 = rm.stats()["nFullRuns"]
        tm = time.perf_counter()
        rm.step(args.sustained_steps)
        ctx.sync()
        elm = time.perf_counter() - tm
        msm, lam = ktm.get("contacts", (float("nan"), 0))
        mtimed = b.n_par * args.steps / elt
        mval = b.n_par * args.sustained_steps / elm
        mixed = {"value": mtimed, "ms_per_step": elt / args.steps * 1e3, "steps": args.steps, "speedup_vs_fp64": mtimed / value,
                 "sustained": {"value": mval, "ms_per_step": elm / args.sustained_steps * 1e3, "steps": args.sustained_steps,
                               "full_collider_runs": rm.stats()["nFullRuns"] - f0, "speedup_vs_fp64": mval / sus["value"]},
                 "contacts_avg_ms": msm / max(lam, 1), "contacts_real": m_real, "contacts_processed": m_proc,
                 "kernels": {k: {"avg_ms": v[0] / max(v[1], 1), "launches": v[1]} for k, v in ktm.items()}}
        m_bytes = (BYTES_REAL["mixed"] * mixed["contacts_real"]
                   + BYTES_POT["mixed"] * max(mixed["contacts_processed"] - mixed["contacts_real"], 0))
        mixed["contacts_roofline"] = {"bound": "hbm", "algorithmic_bytes_per_launch": m_bytes,
                                      "achieved": m_bytes / (mixed["contacts_avg_ms"] * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s"}
        mixed["contacts_roofline"]["frac"] = mixed["contacts_roofline"]["achieved"] / HBM_PEAK_GBS
        rm.close()
        from woodem_amd.precision import mixed_error_report
        mixed["errors_vs_fp64"] = mixed_error_report(ctx)
        log("mixed precision: %s" % json.dumps(mixed))

    copy = None
    try:
        if not args.no_copy:
            copy = hbm_copy_rate()
            log("hbm copy: %s" % json.dumps(copy))
    except Exception as e:  # reported, not fatal
        log("hbm copy failed: %r" % (e,))

    configs = {}
    if not args.no_legs:
        configs["cfg1"] = bench_cfg1(ctx, cpu, cpu_on=not args.no_cpu)
        log("cfg1: %s" % json.dumps(configs["cfg1"]))
        bargs = argparse.Namespace(**vars(args))
        bargs.scenes, bargs.steps, bargs.warmup, bargs.autotune = 512, args.batch_steps, 20, 0
        bargs.seq_scenes, bargs.seq_steps, bargs.latency_steps, bargs.timing_steps = 16, 200, 20, 50
        ctx.set_tuning(window=16, graphs=not args.no_graphs, cell_factor=1.0)
        bt = bench_batch(bargs, ctx)
        configs["cfg4"] = {"workload": bt["config"]["workload"] + ", %d timed steps" % bt["steps"], "value": bt["value"],
                           "unit": bt["unit"], "ms_per_step": bt["ms_per_step"], "sequential": bt["sequential"],
                           "latency_ms": bt["latency_ms"], "roofline": bt["roofline"],
                           "launch": bt["config"]["launch"], "scene_full_runs": bt["config"]["scene_full_runs_in_timed_region"]}
        log("cfg4: %s" % json.dumps(configs["cfg4"]))
        sargs = argparse.Namespace(**vars(args))
        sv = bench_serve(sargs, ctx)
        configs["cfg5"] = {"workload": sv["config"]["workload"], "value": sv["value"], "unit": sv["unit"],
                           "latency_s": sv["latency_s"], "step_latency_ms": sv["step_latency_ms"],
                           "admission_wait_s": sv["admission_wait_s"], "max_resident": sv["max_resident"], "wall_s": sv["wall_s"],
                           "sequential": sv.get("sequential"), "rejected": sv["config"].get("rejected"),
                           "evicted": sv["config"].get("evicted")}
        log("cfg5: %s" % json.dumps(configs["cfg5"]))
        ctx.set_tuning(window=args.window, graphs=not args.no_graphs, cell_factor=args.cell_factor)

    cpub = None
    if not args.no_cpu:
        try:
            cpub = cpu_baseline(ctx, args.cpu_sample, args.cpu_seconds, cpu)
            if gen is not None:
                gen["cpu_baseline"] = cpu_make_cloud(args.cloud_cpu_sample)
            log("cpu baseline: %s" % json.dumps(cpub))
        except Exception as e:  # reported, not fatal for the GPU figure
            log("cpu baseline failed: %r" % (e,))

    return {
        "metric": "particle-steps/s on 1 MI355X (FP64 vs mixed, vs host OpenMP); p50/p99 step latency",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64-state/f32-contact",
        "data": "synthetic (seeded generator, relaxed on device)",
        "config": {"workload": "cfg2: periodic Hertz, %d spheres, phi=0.60, lognormal r (median 1 mm, sigma .3), %d timed steps "
                               "after %d warm-up steps" % (b.n_par, args.steps, args.warmup),
                   "n_particles": b.n_par, "precision": args.precision, "dt": b.meta["dt"], "broadphase": args.broadphase,
                   "contacts_potential": n_con, "contacts_real": n_real, "contacts_processed": n_active,
                   "full_collider_runs_in_timed_region": st1["nFullRuns"] - st0["nFullRuns"], "launch": launch_config(ctx, None)},
        "sustained": sus,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)), "samples": args.latency_steps},
        "roofline": {"bound": "hbm", "kernel": "k_contacts", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "algorithmic_bytes_per_launch": alg_bytes, "avg_launch_ms": avg_ms, "launches": launches,
                     "contacts_real": n_real, "contacts_processed": n_active,
                     "traffic_GBps": (traffic / (avg_ms * 1e-3) / 1e9) if traffic else None, "device_copy": copy},
        "kernels": kernels,
        "mixed_precision": mixed,
        "configs": configs,
        "scene_generation": gen,
        "scene_io": io,
        "cpu_baseline": cpub,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 1000; 5000 for --workload batch)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--relax", type=int, default=500)
    ap.add_argument("--latency-steps", type=int, default=200)
    ap.add_argument("--timing-steps", type=int, default=300,
                    help="steps timed kernel by kernel with HIP events (one at a time; spans collider intervals)")
    ap.add_argument("--cpu-sample", type=int, default=100_000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-io", action="store_true", help="skip the checkpoint/VTK and insertion measurements")
    ap.add_argument("--cloud-n", type=int, default=100_000, help="spheres of the timed device makeCloud (0: skip)")
    ap.add_argument("--cloud-cpu-sample", type=int, default=20_000, help="spheres of the oracle makeCloud sample")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision comparison run")
    ap.add_argument("--workload", choices=["periodic", "batch", "serve"], default="periodic")
    ap.add_argument("--requests", type=int, default=100, help="serve workload: number of requests")
    ap.add_argument("--rate", type=float, default=20., help="serve workload: mean arrivals per second")
    ap.add_argument("--req-steps", type=int, default=10000, help="serve workload: steps per request")
    ap.add_argument("--chunk", type=int, default=100, help="serve workload: steps per batched call")
    ap.add_argument("--hbm-budget", type=float, default=200e9, help="serve workload: bytes the resident scenes may use")
    ap.add_argument("--seed", type=int, default=5, help="serve workload: arrival and scene seed")
    ap.add_argument("--scenes", type=int, default=512, help="batch workload: number of scenes")
    ap.add_argument("--seq-scenes", type=int, default=32, help="batch workload: scenes in the sequential sample")
    ap.add_argument("--seq-steps", type=int, default=300, help="batch workload: steps per scene in the sequential sample")
    ap.add_argument("--relax-box", type=int, default=200, help="batch workload: overlap relaxation sweeps per scene")
    ap.add_argument("--window", type=int, default=16, help="steps launched together as one graph (1 = step by step)")
    ap.add_argument("--no-graphs", action="store_true", help="launch windows without capturing graphs")
    ap.add_argument("--cell-factor", type=float, default=1.0, help="collider cell width over the largest sphere bound")
    ap.add_argument("--no-copy", action="store_true", help="skip the plain device copy that measures the HBM rate (torch)")
    ap.add_argument("--broadphase", choices=["grid", "sweep"], default="grid",
                    help="collider broadphase: multi-level grid (default) or sort and sweep (the same contact set)")
    ap.add_argument("--autotune", type=int, default=None,
                    help="batch workload: pick window and cell factor by the autotuner (default 1 for batch, 0 otherwise)")
    ap.add_argument("--tune-steps", type=int, default=100, help="autotuner: timed steps per candidate and round")
    ap.add_argument("--seq-requests", type=int, default=8, help="serve workload: requests in the sequential sample")
    ap.add_argument("--sustained-steps", type=int, default=1000,
                    help="default line: steps of the sustained rate (windows with their collider full runs; 0: skip)")
    ap.add_argument("--batch-steps", type=int, default=500, help="default line: timed steps of the cfg4 leg")
    ap.add_argument("--no-legs", action="store_true", help="default line: skip the cfg1/cfg3, cfg4 and cfg5 legs")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain device copy of 1 GiB (5 times), for scaling FETCH_SIZE/WRITE_SIZE counters")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 5000 if args.workload == "batch" else 1000
    if args.autotune is None:
        args.autotune = 1 if args.workload == "batch" else 0
    if args.calibrate_copy:
        # kernels of known HBM traffic for scaling the FETCH_SIZE/WRITE_SIZE counters: a 1 GiB
        # copy (1 GiB read, 1 GiB written), a 1 GiB sum (read only) and a gather of 2^22 random
        # 64-byte rows of a 1 GiB table (256 MiB read in single lines, 256 MiB written)
        import torch
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        s = None
        for _ in range(5):
            s = a.sum()
        rows = a.view(-1, 8)
        idx = torch.randint(0, rows.shape[0], (1 << 22,), device="cuda", generator=torch.Generator("cuda").manual_seed(1))
        for _ in range(5):
            g = rows.index_select(0, idx)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 8, "copy_bytes_written": a.numel() * 8, "sum_bytes_read": a.numel() * 8,
                          "gather_bytes_read": g.numel() * 8, "gather_bytes_written": g.numel() * 8, "sum": float(s)}))
        return
    if args.gpus != 1:
        print("bench.py: this build runs the timestep on one GPU only; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    from woodem_amd import runner, scenes
    from woodem_amd._native import PREC_FP64, PREC_MIXED, Context

    ctx = Context(0)
    ctx.set_tuning(window=args.window, graphs=not args.no_graphs, cell_factor=args.cell_factor)
    ctx.set_broadphase(1 if args.broadphase == "sweep" else 0)
    if args.workload == "batch":
        print(json.dumps(bench_batch(args, ctx)), flush=True)
        return
    if args.workload == "serve":
        print(json.dumps(bench_serve(args, ctx)), flush=True)
        return
    print(json.dumps(bench_periodic(args, ctx)), flush=True)


if __name__ == "__main__":
    main()
