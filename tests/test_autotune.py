"""The launch autotuner's search logic (on a stand-in context: no GPU)."""
from edge_matching_puzzle_amd import autotune


class FakeContext:
    """Rate peaks at waves_per_cu 16, steps_per_round 64; counts calls."""

    def __init__(self):
        self.calls = 0

    def device_info(self):
        return {"compute_units": 4}

    def enumerate(self, puzzles, **kw):
        self.calls += 1
        score = 100 - abs(kw["waves_per_cu"] - 16) - abs(kw["steps_per_round"] - 64) / 8
        return [{"n_nodes": int(score * 1000)}], {}


def test_grid_covers_the_space():
    space = {"waves_per_cu": [8, 16], "waves_per_block": [4], "steps_per_round": [32, 64],
             "target_per_lane": [4], "step_budget": [1 << 16]}
    settings = list(autotune.grid(space))
    assert len(settings) == 4
    assert {(s["waves_per_cu"], s["steps_per_round"]) for s in settings} == {(8, 32), (8, 64), (16, 32), (16, 64)}


def test_coordinate_search_finds_the_best_axis_values(monkeypatch):
    ctx = FakeContext()
    clock = iter(range(10000))
    monkeypatch.setattr(autotune.time, "perf_counter", lambda: next(clock) * 1e-3)
    best, history = autotune.tune(ctx, [None], node_limit=1, budget_s=1e9)
    assert best["waves_per_cu"] == 16 and best["steps_per_round"] == 64
    assert len(history) == ctx.calls
    # coordinate search: each trial changes one axis of the best setting so far
    for (a, _), (b, _) in zip(history, history[1:]):
        assert sum(a[k] != b[k] for k in a) <= 2
