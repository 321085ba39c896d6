"""The CPU oracle against the reference's own fixtures (tests/golden/)."""
import os

import pytest

import oracle_bindings as o

SIZES = ("3_3", "4_4", "5_5")
# tests/new_strategy_{3_3,4_4,5_5}/test.info:19-20,33-34,51-52
TOTALS = {"3_3": (37, 2), "4_4": (3284, 12), "5_5": (37938, 1)}


@pytest.mark.parametrize("name", SIZES)
def test_node_and_solution_totals(example_puzzles, name):
    w, h, c = example_puzzles[name]
    r = o.enumerate_(w, h, c)
    assert (r["n_nodes"], r["n_solutions"]) == TOTALS[name]


@pytest.mark.parametrize("name", SIZES)
def test_results_bin_byte_identical(example_puzzles, golden_dir, name):
    w, h, c = example_puzzles[name]
    with open(os.path.join(golden_dir, "new_strategy_%s__results.bin" % name), "rb") as f:
        want = f.read()
    assert o.results_bin(w, h, c) == want


@pytest.mark.parametrize("name", SIZES)
def test_new_max_trace(example_puzzles, expected_stdout, name):
    """The MAX_DISPLAY lines, in order, as the reference test greps them."""
    w, h, c = example_puzzles[name]
    r = o.enumerate_(w, h, c, trace=True)
    lines = r["trace"].splitlines() + ["Total situations explored : %d" % r["n_nodes"],
                                       "Nb solutions : %d" % r["n_solutions"]]
    want = expected_stdout["new_strategy_" + name]["strings"]
    i = 0
    for line in lines:
        if i < len(want) and want[i] in line:
            i += 1
    assert i == len(want)


def _parse_stats(path):
    invalid, intermediate, section = {}, {}, None
    for line in open(path):
        line = line.strip()
        if line.startswith("Nunber of invalid"):
            section = invalid
        elif line.startswith("Nunber of intermediate"):
            section = intermediate
        elif line.startswith("Total"):
            section = None
        elif section is not None and ":" in line:
            level, value = line.split(":")
            section[int(level)] = int(value)
    return invalid, intermediate


@pytest.mark.parametrize("name", SIZES)
def test_per_depth_histogram(example_puzzles, golden_dir, name):
    """compute_stats tables: level L holds the situations with L pieces."""
    w, h, c = example_puzzles[name]
    r = o.enumerate_(w, h, c)
    invalid, intermediate = _parse_stats(os.path.join(golden_dir, "compute_stats_%s__stats_%s.txt" % (name, name)))
    P = w * h
    for level in range(1, P):
        k = level - 1
        assert invalid[level] == r["depth_dead"][k]
        assert intermediate[level] == r["depth_nodes"][k] - r["depth_dead"][k]
    assert r["depth_nodes"][P - 1] == r["n_solutions"]


def test_display_solutions_4_4_strings(example_puzzles, expected_stdout):
    """Same 12 solutions as the FSM path lists (tests/display_solutions_4_4)."""
    w, h, c = example_puzzles["4_4"]
    r = o.enumerate_(w, h, c)
    got = sorted(o.situation_string(b, w * h) for b in r["boards"])
    want = sorted(s.split('"')[1] for s in expected_stdout["display_solutions_4_4"]["strings"]
                  if s.startswith("Final situation"))
    assert got == want


def _rotate(board, w, h):
    """Quarter turn clockwise of a row-major board (square boards)."""
    out = [0xFFFF] * (w * h)
    for y in range(h):
        for x in range(w):
            v = int(board[y * w + x])
            nx, ny = w - 1 - y, x
            out[ny * w + nx] = (v & ~3) | ((v + 3) & 3) if v != 0xFFFF else v
    return out


def test_rotations_give_profile_solutions(example_puzzles, golden_dir):
    """tests/profile_3_3/references/solutions.txt: all 8 unbroken solutions."""
    w, h, c = example_puzzles["3_3"]
    r = o.enumerate_(w, h, c)
    boards = set()
    for b in r["boards"]:
        cur = list(int(v) for v in b)
        for _ in range(4):
            boards.add(o.situation_string(cur, 9))
            cur = _rotate(cur, w, h)
    want = [l.strip() for l in open(os.path.join(golden_dir, "profile_3_3__solutions.txt")) if l.strip()]
    assert sorted(boards) == sorted(want)


@pytest.mark.parametrize("name", SIZES)
def test_glutton_max_known_answers(example_puzzles, expected_stdout, name):
    """tests/CUDA_glutton_max_{3_3,4_4,5_5}: first solution, tie-break exact."""
    w, h, c = example_puzzles[name]
    r = o.glutton_max(w, h, c)
    assert r["solved"]
    assert o.situation_string(r["board"], w * h) in expected_stdout["CUDA_glutton_max_" + name]["strings"]


def test_spiral_order_matches_results_bin_header(golden_dir):
    import struct
    for name, n in (("3_3", 3), ("4_4", 4), ("5_5", 5)):
        data = open(os.path.join(golden_dir, "new_strategy_%s__results.bin" % name), "rb").read()
        xy = struct.unpack("<%dI" % (2 * n * n), data[16:16 + 8 * n * n])
        assert o.order(n, n) == [(xy[2 * i], xy[2 * i + 1]) for i in range(n * n)]


def test_node_limit_reports_partial():
    w, h, c = o.parse_puzzle(os.path.join(os.path.dirname(__file__), "golden", "puzzles", "5_5.txt"))
    r = o.enumerate_(w, h, c, node_limit=1000)
    assert r["status"] == o.LIMIT_REACHED and r["n_nodes"] == 1000


@pytest.mark.parametrize("w,B,I,seed,depth", [(5, 3, 4, 1, 6), (5, 2, 2, 4, 4), (6, 3, 6, 1, 8)])
def test_prefix_subtrees_union_is_the_enumeration(w, B, I, seed, depth):
    """The cut used to build the 8x8 golden in parallel: nodes above the cut
    plus every prefix subtree equal the uncut enumeration, solution for
    solution."""
    import edge_matching_puzzle_amd as emp
    c, _ = emp.generate(seed, w, w, B, I)
    full = o.enumerate_(w, w, c, max_solutions=1 << 18)
    pre, nodes, sols = o.prefixes(w, w, c, depth)
    boards = []
    for p in pre:
        r = o.subtree(w, w, c, p, max_solutions=1 << 18)
        nodes += r["n_nodes"]
        sols += r["n_solutions"]
        boards += [tuple(b) for b in r["boards"].tolist()]
    assert nodes == full["n_nodes"] and sols == full["n_solutions"]
    assert sorted(boards) == sorted(tuple(b) for b in full["boards"].tolist())
