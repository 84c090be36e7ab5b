"""The launch-shape autotuner's search on a modelled context (CPU only): it times every
candidate on the scenes themselves, keeps the fastest window, then the fastest cell factor
at that window, sets the winner, and every step it took counts."""
from woodem_amd.autotune import autotune


class Clock:
    def __init__(self):
        self.t = 0.

    def __call__(self):
        return self.t


class ModelCtx:
    """ms per step = a bowl in (window, cell factor) with its minimum at (16, 1.3)."""

    def __init__(self, clock):
        self.clock = clock
        self.window, self.graphs, self.cell_factor = 16, 1, 1.
        self.steps = 0
        self.sets = []

    def set_tuning(self, window=None, graphs=None, cell_factor=None):
        if window is not None:
            self.window = window
        if graphs is not None:
            self.graphs = graphs
        if cell_factor is not None:
            self.cell_factor = cell_factor
        self.sets.append((self.window, self.graphs, self.cell_factor))

    def tuning(self):
        return dict(window=self.window, graphs=self.graphs, cell_factor=self.cell_factor)

    def step(self, sids, n):
        ms = 1. + abs(self.window - 16) * .01 + abs(self.cell_factor - 1.3) * .2
        self.clock.t += n * ms * 1e-3
        self.steps += n

    def sync(self):
        pass


def test_autotune_picks_the_fastest_and_counts_its_steps():
    clk = Clock()
    ctx = ModelCtx(clk)
    ctx.cell_factor = 1.7  # the window search runs at the current factor
    r = autotune(ctx, [0, 1, 2], steps=10, warm=2, windows=(1, 8, 16, 64), cell_factors=(1., 1.3, 1.7), rounds=2, clock=clk)
    assert (r["window"], r["cell_factor"]) == (16, 1.3)
    assert ctx.tuning()["window"] == 16 and ctx.tuning()["cell_factor"] == 1.3 and ctx.tuning()["graphs"] == 1
    assert len(r["trials"]) == 2 * (4 + 3)
    assert r["steps"] == ctx.steps == len(r["trials"]) * 12
    # window trials all ran at the starting cell factor, factor trials at the chosen window
    assert all(t["cell_factor"] == 1.7 for t in r["trials"][:8])
    assert all(t["window"] == 16 for t in r["trials"][8:])
