"""Autotuner for the launch shape of wdm_step.

Two knobs decide how a batch of scenes is launched, and neither changes a result
(tests/test_gpu_window.py checks bit equality): the window (steps launched together as one
captured HIP graph between two looks of the host at the Verlet triggers) and the collider's
grid cell width (bin size) as a factor over the largest sphere bound.  Long windows save
host round trips, but a scene whose trigger fires waits for the rest of the window; coarse
cells shorten the grid passes of small scenes but put more candidate pairs in each cell.
The best setting depends on the batch, so it is measured on the batch itself: each
candidate steps the scenes for real (the steps count, they are not thrown away) and the
fastest per step wins.  The block size stays fixed at 256 threads: the kernels' per-block
energy partial sums define the summation order the parity tests pin.
"""
import time


def autotune(ctx, sids, steps=40, warm=8, windows=(1, 4, 8, 16, 32, 64), cell_factors=(1., 1.3, 1.7), rounds=2,
             clock=time.perf_counter, log=None):
    """Pick (window, cell_factor) for stepping `sids` together on context `ctx`; sets the
    winner on the context and returns {"window", "cell_factor", "graphs", "trials", "steps"}.

    Coordinate search: the window first (at the current cell factor), then the cell factor
    at the best window.  Candidates are timed in `rounds` interleaved rounds of `warm`
    untimed plus `steps` timed steps each (so a slow drift of the scenes does not favour
    the first candidate); a candidate's time per step is its best round."""
    sids = list(sids)
    total = [0]
    trials = []

    def timed(window, factor):
        ctx.set_tuning(window=window, graphs=1, cell_factor=factor)
        ctx.step(sids, warm)  # graph capture, buffer growth
        ctx.sync()
        t = clock()
        ctx.step(sids, steps)
        ctx.sync()
        total[0] += warm + steps
        return (clock() - t) / steps

    def search(cands):
        best = {c: float("inf") for c in cands}
        for r in range(rounds):
            for c in (cands if r % 2 == 0 else cands[::-1]):
                ms = timed(*c) * 1e3
                best[c] = min(best[c], ms)
                trials.append(dict(window=c[0], cell_factor=c[1], round=r, ms_per_step=ms))
                if log:
                    log("autotune: window %d, cell factor %.2f: %.4f ms/step" % (c[0], c[1], ms))
        return min(cands, key=lambda c: best[c])

    f0 = float(ctx.tuning()["cell_factor"])
    w = search([(int(x), f0) for x in windows])[0]
    f = search([(w, float(x)) for x in cell_factors])[1]
    ctx.set_tuning(window=w, graphs=1, cell_factor=f)
    return dict(window=w, cell_factor=f, graphs=1, trials=trials, steps=total[0])
