"""GPU CUDA-backtracker search (first solution + selection trace) against the
oracle's restatement and the reference's own trace (tests_CUDA/backtracker_*)."""
import pytest

import edge_matching_puzzle_amd as emp
import oracle_bindings as o

pytestmark = pytest.mark.gpu


def _lines(trace):
    return ["Thread %d : selected variable %d at step %d" % tuple(int(v) for v in t) for t in trace]


@pytest.mark.parametrize("name", ["3_3", "4_4", "5_5"])
def test_reference_trace_and_solution(gpu_context, example_puzzles, expected_stdout, name):
    w, h, c = example_puzzles[name]
    print("case: backtrack", name, flush=True)
    r = gpu_context.backtrack_trace(w, h, c)
    want = expected_stdout["backtracker_" + name]["strings"]
    assert r["found"]
    assert emp.situation_string(w, h, r["board"]) == want[-1]
    if name == "3_3":
        assert _lines(r["trace"]) == want[:-1]      # the 12 selections, in order
    ref = o.backtracker(w, h, c)
    assert r["trace_length"] == ref["trace_length"]
    assert (r["trace"] == ref["trace"]).all()


@pytest.mark.parametrize("w,B,I,seed,order", [(5, 3, 4, 1, "basic"), (5, 3, 4, 2, "spiral"), (6, 3, 6, 3, "basic"),
                                              (4, 2, 2, 4, "basic"), (6, 4, 5, 5, "spiral")])
def test_synthetic_against_oracle(gpu_context, w, B, I, seed, order):
    c, _ = emp.generate(seed, w, w, B, I)
    print("case: backtrack synthetic", w, seed, order, flush=True)
    r = gpu_context.backtrack_trace(w, w, c, order=order, trace_capacity=1 << 18)
    ref = o.backtracker(w, w, c, order=o.BASIC if order == "basic" else o.SPIRAL, trace_capacity=1 << 18)
    assert r["found"] == ref["found"]
    assert r["trace_length"] == ref["trace_length"]
    assert (r["trace"] == ref["trace"]).all()
    assert list(r["board"]) == list(ref["board"])


@pytest.mark.parametrize("w,h,B,I,seed,order", [(6, 4, 2, 3, 1, "spiral"), (5, 3, 1, 2, 2, "basic"),
                                                (4, 6, 2, 3, 3, "basic"), (7, 4, 2, 4, 4, "spiral"),
                                                (3, 7, 2, 3, 5, "basic"), (5, 4, 1, 3, 6, "spiral")])
def test_rectangles_against_oracle(gpu_context, w, h, B, I, seed, order):
    """Wide and tall few-colour boards; 5x3 with one border colour has no
    solution, so its trace is the whole exhausted search (121,556 entries)."""
    c, _ = emp.generate(seed, w, h, B, I)
    print("case: backtrack rectangle", w, h, B, I, seed, order, flush=True)
    r = gpu_context.backtrack_trace(w, h, c, order=order, trace_capacity=1 << 18)
    ref = o.backtracker(w, h, c, order=o.BASIC if order == "basic" else o.SPIRAL, trace_capacity=1 << 18)
    assert r["found"] == ref["found"]
    assert r["trace_length"] == ref["trace_length"]
    assert (r["trace"] == ref["trace"]).all()
    assert list(r["board"]) == list(ref["board"])


def test_selection_limit(gpu_context):
    c, _ = emp.generate(1, 6, 6, 3, 6)
    r = gpu_context.backtrack_trace(6, 6, c, max_selections=10)
    assert r["trace_length"] == 10 and not r["found"]
