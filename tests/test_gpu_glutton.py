"""GPU glutton_max (through the C ABI) against the CPU oracle's restatement.

Bar: same status, same max depth, same number of pushes and the same board
(the solution, or the deepest situation) for every search.
"""
import numpy as np
import pytest

import edge_matching_puzzle_amd as emp
import oracle_bindings as o

pytestmark = pytest.mark.gpu


def _unpack(row):
    return [(int(m) & 0xFF, (int(m) >> 8) & 0xFF, (int(m) >> 16) & 3) for m in row]


def _oracle(w, h, c, row, budget):
    moves = [v for t in _unpack(row) for v in t]
    return o.glutton_max(w, h, c, prefix=moves, push_limit=budget)


def _check(gpu, board, ref):
    assert gpu["max_depth"] == ref["max_depth"]
    assert gpu["pushes"] == ref["pushes"]
    assert (gpu["status"] == emp.GLUTTON_SOLVED) == ref["solved"]
    assert (gpu["status"] == emp.GLUTTON_BUDGET) == (ref["status"] == o.LIMIT_REACHED)
    assert list(board) == list(ref["board"])


@pytest.mark.parametrize("name", ["3_3", "4_4", "5_5"])
def test_known_answers(gpu_context, example_puzzles, expected_stdout, name):
    """tests/CUDA_glutton_max_{3_3,4_4,5_5}: the first solution, tie-break exact."""
    w, h, c = example_puzzles[name]
    print("case: glutton example", name, flush=True)
    (r,), boards, st = gpu_context.glutton_max(w, h, c)
    assert r["status"] == emp.GLUTTON_SOLVED
    assert emp.situation_string(w, h, boards[0]) in expected_stdout["CUDA_glutton_max_" + name]["strings"]
    _check(r, boards[0], o.glutton_max(w, h, c))


def test_synthetic_empty_start(gpu_context):
    for seed in range(6):
        w = h = 5
        c, _ = emp.generate(seed, w, h, 3, 4)
        print("case: glutton 5x5 seed", seed, flush=True)
        (r,), boards, _ = gpu_context.glutton_max(w, h, c, push_budget=400)
        _check(r, boards[0], o.glutton_max(w, h, c, push_limit=400))


@pytest.mark.parametrize("w,B,I,seed,length,budget", [(6, 3, 6, 1, 6, 60), (7, 4, 7, 2, 8, 40),
                                                      (8, 4, 8, 3, 10, 25), (6, 2, 3, 4, 5, 80)])
def test_prefixed_searches(gpu_context, w, B, I, seed, length, budget):
    """Several prefixes in one launch, each played before the search."""
    c, _ = emp.generate(seed, w, w, B, I)
    rows = emp.spiral_prefixes(w, w, c, length, 12, seed)
    assert len(rows) >= 4
    print("case: glutton prefixed", w, B, I, seed, len(rows), flush=True)
    res, boards, st = gpu_context.glutton_max(w, w, c, prefixes=rows, push_budget=budget)
    for i, row in enumerate(rows):
        _check(res[i], boards[i], _oracle(w, w, c, row, budget))


@pytest.mark.parametrize("w,h,B,I,seed,budget", [(6, 4, 2, 3, 1, 300), (7, 5, 2, 3, 2, 300),
                                                 (8, 5, 3, 4, 3, 200), (5, 3, 1, 2, 4, 300),
                                                 (9, 6, 2, 4, 5, 150), (4, 7, 2, 3, 6, 300),
                                                 (6, 6, 1, 2, 7, 300), (10, 4, 2, 3, 8, 200)])
def test_rectangles_and_few_colours(gpu_context, w, h, B, I, seed, budget):
    """Wide and tall boards with few colours: many identical pieces, so the
    most-constrained choice meets many ties.  Six solve, two stop at the
    push budget."""
    c, _ = emp.generate(seed, w, h, B, I)
    print("case: glutton rectangle", w, h, B, I, seed, budget, flush=True)
    (r,), boards, _ = gpu_context.glutton_max(w, h, c, push_budget=budget)
    _check(r, boards[0], o.glutton_max(w, h, c, push_limit=budget))


def test_bad_prefix_is_reported(gpu_context):
    w = h = 5
    c, _ = emp.generate(3, w, h, 3, 4)
    row = emp.spiral_prefixes(w, h, c, 4, 1, 3)[0].copy()
    row[1] = row[0]                       # same cell twice
    (r,), _, _ = gpu_context.glutton_max(w, h, c, prefixes=row[None, :], push_budget=5)
    assert r["status"] == emp.GLUTTON_BAD_PREFIX


def test_c3_shape_against_golden(gpu_context, golden_dir):
    """16x16, 5 border / 17 interior colours, centre hint + 24 spiral moves,
    two pushes; expected results from tests/golden/make_glutton_16x16_golden.py
    (the oracle needs about half a minute per push at this size)."""
    import json
    import os
    g = json.load(open(os.path.join(golden_dir, "glutton_16x16_seed3.json")))
    w = h = 16
    c, _ = emp.generate(g["seed"], w, h, g["border_colours"], g["interior_colours"])
    rows = np.array([s["prefix"] for s in g["searches"]], dtype=np.uint32)
    print("case: glutton 16x16 golden", flush=True)
    res, boards, st = gpu_context.glutton_max(w, h, c, prefixes=rows, push_budget=g["push_budget"])
    for i, s in enumerate(g["searches"]):
        ref = {"max_depth": s["max_depth"], "pushes": s["pushes"], "solved": s["solved"],
               "status": o.LIMIT_REACHED if s["budget_reached"] else o.OK, "board": s["board"]}
        _check(res[i], boards[i], ref)
