"""The oracle's CUDA-backtracker restatement against tests_CUDA/backtracker_*."""
import pytest

import oracle_bindings as o


@pytest.mark.parametrize("name", ["3_3", "4_4", "5_5"])
def test_trace_and_first_solution(example_puzzles, expected_stdout, name):
    w, h, c = example_puzzles[name]
    r = o.backtracker(w, h, c)
    want = expected_stdout["backtracker_" + name]["strings"]
    assert r["found"] and o.situation_string(r["board"], w * h) == want[-1]
    lines = ["Thread %d : selected variable %d at step %d" % tuple(int(v) for v in t) for t in r["trace"]]
    for s in want[:-1]:
        assert s in lines
    if name == "3_3":
        assert lines == want[:-1]


def test_first_solution_is_first_in_basic_enumeration(example_puzzles):
    """Orientation-major candidate order is a different DFS order than the
    enumeration's, but the solution found must be one of the enumeration's."""
    for name in ("3_3", "4_4", "5_5"):
        w, h, c = example_puzzles[name]
        r = o.backtracker(w, h, c)
        sols = {tuple(int(v) for v in b) for b in o.enumerate_(w, h, c, order=o.BASIC)["boards"]}
        assert tuple(int(v) for v in r["board"]) in sols
