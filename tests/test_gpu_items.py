"""Ordered results, resumed searches and resumable cuts of the GPU
enumeration (emp_hip_enumerate_items, through the C ABI), and the
new_strategy command that writes results.bin from them.

Bars:
  - results.bin of the GPU enumeration byte-identical to the reference's
    fixtures (tests/new_strategy_*/references/results.bin);
  - node indices equal to the oracle's sequential search, with work items
    split and requeued many times (the order is rebuilt from the items);
  - a search resumed from a situation equal to the oracle's restatement of
    the reference's resume (nodes, solutions, their order and indices);
  - a run cut by its node limit and resumed from the work it returns, until
    nothing is left, adds up to the uninterrupted run (nodes and solutions).
"""
import os
import subprocess

import numpy as np
import pytest

import edge_matching_puzzle_amd as emp
import oracle_bindings as o

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "edge_matching_puzzle_amd", "edge_matching_puzzle")
PUZZLES = os.path.join(ROOT, "tests", "golden", "puzzles")


def _ordered_same(gpu, ref, nodes=None):
    assert gpu["complete"]
    assert gpu["n_nodes"] == (ref["n_nodes"] if nodes is None else nodes)
    assert gpu["n_solutions"] == ref["n_solutions"]
    assert [tuple(b) for b in gpu["boards"]] == [tuple(b) for b in ref["boards"]]
    assert [int(v) for v in gpu["node_index"]] == [int(v) for v in ref["solution_nodes"]]


@pytest.mark.parametrize("name", ["3_3", "4_4", "5_5"])
def test_results_bin_is_the_reference_fixture(gpu_context, example_puzzles, golden_dir, name):
    w, h, c = example_puzzles[name]
    print("case: ordered results", name, flush=True)
    r = gpu_context.enumerate_items(w, h, c, ordered=True)
    data = emp.results_bin_write(w, h, r["boards"], r["node_index"], r["n_nodes"])
    assert data == open(os.path.join(golden_dir, "new_strategy_%s__results.bin" % name), "rb").read()


@pytest.mark.parametrize("w,h,B,I,seed,budget", [(5, 5, 3, 4, 11, 0), (5, 5, 2, 3, 12, 64), (6, 6, 3, 6, 1, 64),
                                                 (6, 5, 2, 4, 13, 256), (6, 4, 2, 2, 14, 32)])
def test_node_indices_match_the_sequential_search(gpu_context, w, h, B, I, seed, budget):
    """Few colours (identical pieces) and tiny step budgets: items split and
    requeue thousands of times, and the order is still the oracle's."""
    c, _ = emp.generate(seed, w, h, B, I)
    ref = o.enumerate_(w, h, c, max_solutions=1 << 14)
    print("case: node indices", w, h, B, I, seed, budget, ref["n_nodes"], flush=True)
    r = gpu_context.enumerate_items(w, h, c, ordered=True, max_solutions=1 << 14, step_budget=budget)
    if budget:
        assert r["stats"]["splits"] > 100
    _ordered_same(r, ref)


def _situation(w, h, board, k):
    xy = emp.order(w, h)
    s = np.full(w * h, 0xFFFF, dtype=np.uint16)
    for x, y in xy[:k]:
        s[y * w + x] = board[y * w + x]
    return s


@pytest.mark.parametrize("name", ["3_3", "4_4", "5_5"])
def test_resume_from_a_situation_matches_the_oracle(gpu_context, example_puzzles, name):
    w, h, c = example_puzzles[name]
    full = o.enumerate_(w, h, c)
    P = w * h
    for board in full["boards"][:4]:
        for k in sorted({1, 2, 4, P // 2, P - 1}):
            sit = _situation(w, h, board, k)
            items, moves = emp.resume_item(w, h, c, sit)
            assert len(items) == (0 if k == 1 else 1)
            print("case: resume", name, k, flush=True)
            r = gpu_context.enumerate_items(w, h, c, items=items, moves=moves, ordered=True)
            _ordered_same(r, o.resume(w, h, c, sit))


@pytest.mark.parametrize("seed", [2, 3])
def test_resume_on_synthetic_puzzles(gpu_context, seed):
    """6x6, 3 border / 6 interior colours (2.5e7 nodes, 2-4 solutions)."""
    w = h = 6
    c, planted = emp.generate(seed, w, h, 3, 6)
    full = o.enumerate_(w, h, c)
    for board in list(full["boards"][:2]) or [planted]:
        for k in (3, 17):
            sit = _situation(w, h, board, k)
            items, moves = emp.resume_item(w, h, c, sit)
            print("case: resume 6x6", seed, k, flush=True)
            r = gpu_context.enumerate_items(w, h, c, items=items, moves=moves, ordered=True, step_budget=128)
            _ordered_same(r, o.resume(w, h, c, sit))


def test_resume_refuses_a_situation_the_search_never_visits(gpu_context, example_puzzles):
    w, h, c = example_puzzles["4_4"]
    full = o.enumerate_(w, h, c)
    sit = _situation(w, h, full["boards"][0], 5)
    x, y = emp.order(w, h)[2]
    sit[y * w + x] ^= 2
    with pytest.raises(emp.EmpError) as e:
        emp.resume_item(w, h, c, sit)
    assert e.value.status == emp.BAD_ARGUMENT


def _chain(ctx, w, h, c, limit, rest_capacity=1 << 18, **launch):
    """Runs cut at `limit` nodes, each resumed from the rest of the one
    before, until nothing is left: (calls, nodes, solutions, boards)."""
    items = moves = None
    calls = nodes = sols = 0
    boards = []
    while True:
        r = ctx.enumerate_items(w, h, c, items=items, moves=moves, max_solutions=4096, node_limit=limit,
                                rest_capacity=rest_capacity, **launch)
        calls += 1
        nodes += r["n_nodes"]
        sols += r["n_solutions"]
        boards += [tuple(b) for b in r["boards"]]
        items, moves = r["rest_items"], r["rest_moves"]
        assert r["complete"] == (len(items) == 0)
        if not len(items):
            return calls, nodes, sols, boards
        assert calls < 64


@pytest.mark.parametrize("w,h,B,I,seed,parts", [(6, 6, 3, 6, 1, 5), (8, 8, 6, 16, 11, 4), (6, 5, 2, 4, 13, 6)])
def test_cut_runs_resumed_add_up_to_the_whole(gpu_context, w, h, B, I, seed, parts):
    c, _ = emp.generate(seed, w, h, B, I)
    (whole,), _ = gpu_context.enumerate([(w, h, c)], max_solutions=4096)
    limit = max(whole["n_nodes"] // parts, 1)
    print("case: cut and resume", w, h, B, I, seed, whole["n_nodes"], limit, flush=True)
    calls, nodes, sols, boards = _chain(gpu_context, w, h, c, limit, step_budget=1 << 12, target_items=1 << 12,
                                        queue_items=1 << 16)
    assert calls >= 2
    assert nodes == whole["n_nodes"] and sols == whole["n_solutions"]
    assert len(boards) == len(set(boards)) == sols
    assert sorted(boards) == sorted(tuple(b) for b in whole["boards"])
    ref = o.enumerate_(w, h, c, max_solutions=4096)
    assert nodes == ref["n_nodes"] and sorted(boards) == sorted(tuple(b) for b in ref["boards"])


def test_time_cut_rest_resumes_to_the_whole(gpu_context):
    """The same with a time limit (the cut falls wherever the clock says)."""
    w = h = 8
    c, _ = emp.generate(11, w, h, 6, 16)
    (whole,), _ = gpu_context.enumerate([(w, h, c)], max_solutions=4096)
    items = moves = None
    nodes = calls = 0
    boards = []
    print("case: time cut 8x8", flush=True)
    while True:
        r = gpu_context.enumerate_items(w, h, c, items=items, moves=moves, max_solutions=4096, time_limit_ms=2.0,
                                        rest_capacity=1 << 18, step_budget=1 << 10, target_items=1 << 12,
                                        queue_items=1 << 16)
        calls += 1
        nodes += r["n_nodes"]
        boards += [tuple(b) for b in r["boards"]]
        items, moves = r["rest_items"], r["rest_moves"]
        if not len(items):
            break
        assert calls < 400
    assert nodes == whole["n_nodes"] and sorted(boards) == sorted(tuple(b) for b in whole["boards"])


def test_rest_beyond_its_capacity_is_an_error(gpu_context):
    w = h = 8
    c, _ = emp.generate(11, w, h, 6, 16)
    with pytest.raises(emp.EmpError) as e:
        gpu_context.enumerate_items(w, h, c, node_limit=1000, rest_capacity=1, step_budget=64)
    assert e.value.status == emp.OUT_OF_MEMORY


def _cli(*args):
    p = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=120)
    return p.returncode, p.stdout


@pytest.mark.parametrize("name,nodes,sols", [("3_3", 37, 2), ("4_4", 3284, 12), ("5_5", 37938, 1)])
def test_new_strategy_command(golden_dir, expected_stdout, tmp_path, name, nodes, sols):
    """The reference's tests/new_strategy_<name>: its totals on stdout and its
    results.bin, byte for byte, in the dump file."""
    dump = tmp_path / "results.bin"
    print("case: CLI new_strategy", name, flush=True)
    rc, out = _cli(os.path.join(PUZZLES, name + ".txt"), "ressources", "new_strategy", "--enable_gui=false",
                   "--dump_file=%s" % dump)
    assert rc == 0, out
    for s in expected_stdout["new_strategy_" + name]["strings"]:
        if s.startswith("Total") or s.startswith("Nb solutions"):
            assert s in out.splitlines()
    assert "Total situations explored : %d" % nodes in out and "Nb solutions : %d" % sols in out
    assert dump.read_bytes() == open(os.path.join(golden_dir, "new_strategy_%s__results.bin" % name), "rb").read()


def test_new_strategy_command_resumes_an_initial_situation(tmp_path, example_puzzles):
    w, h, c = example_puzzles["4_4"]
    full = o.enumerate_(w, h, c)
    sit = _situation(w, h, full["boards"][5], 7)
    text = emp.situation_string(w, h, sit)
    ref = o.resume(w, h, c, sit)
    dump = tmp_path / "resumed.bin"
    print("case: CLI new_strategy --initial_situation", flush=True)
    rc, out = _cli(os.path.join(PUZZLES, "4_4.txt"), "ressources", "new_strategy", "--initial_situation=" + text,
                   "--dump_file=%s" % dump)
    assert rc == 0, out
    assert 'Setting initial state @ "%s"' % text in out
    assert "Total situations explored : %d" % ref["n_nodes"] in out
    assert "Nb solutions : %d" % ref["n_solutions"] in out
    r = emp.results_bin_read(dump.read_bytes())
    assert [tuple(b) for b in r["boards"]] == [tuple(b) for b in ref["boards"]]
    assert list(r["node_index"]) == list(ref["solution_nodes"]) and r["total_nodes"] == ref["n_nodes"]
