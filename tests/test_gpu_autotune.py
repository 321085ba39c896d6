"""The launch autotuner on the device: every setting it tries enumerates the
batch exactly (each puzzle's node count, solution count and sorted solution
boards equal the oracle's), and the setting it returns is one of the
measured ones with the best rate."""
import numpy as np
import pytest

import edge_matching_puzzle_amd as emp
from edge_matching_puzzle_amd import autotune
import oracle_bindings as o

pytestmark = pytest.mark.gpu


def _batch():
    """Uneven sizes (4..6), 3..5 border and 6..12 interior colours: the C2
    shape at sizes the oracle enumerates in a moment."""
    rng = np.random.default_rng(17)
    puzzles = []
    for s in range(32):
        w = int(rng.integers(4, 7))
        puzzles.append((w, w, emp.generate(5000 + s, w, w, int(rng.integers(3, 6)), int(rng.integers(6, 13)))[0]))
    return puzzles


def test_every_tuned_setting_is_exact(gpu_context):
    puzzles = _batch()
    refs = [o.enumerate_(w, h, c, max_solutions=4096) for w, h, c in puzzles]
    space = {"waves_per_cu": [8, 16], "waves_per_block": [2, 4], "steps_per_round": [16, 32],
             "target_per_lane": [1, 4], "step_budget": [1 << 10, 1 << 16]}
    seen = []

    def verify(setting, res):
        print("case: autotune trial", setting, flush=True)
        seen.append(dict(setting))
        for (w, h, c), r, ref in zip(puzzles, res, refs):
            assert r["complete"]
            assert (r["n_nodes"], r["n_solutions"]) == (ref["n_nodes"], ref["n_solutions"]), setting
            if ref["n_solutions"] <= 4096:
                assert sorted(map(tuple, r["boards"].tolist())) == sorted(map(tuple, ref["boards"].tolist())), setting

    best, history = autotune.tune(gpu_context, puzzles, 0, space=space, budget_s=1e9, verify=verify,
                                  max_solutions=4096, use_graph=True)
    # the middle of each axis first, then each other value of an axis (and,
    # after an improvement, the axis's earlier best again)
    assert len(history) == len(seen)
    assert 1 + sum(len(v) - 1 for v in space.values()) <= len(history) <= 1 + sum(len(v) for v in space.values())
    assert best in [s for s, _ in history]
    assert dict(history[0][0]) == {k: v[len(v) // 2] for k, v in space.items()}
    best_rate = max(r for _, r in history)
    assert any(s == best and r == best_rate for s, r in history)
