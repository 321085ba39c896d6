"""The oracle's restatement of the reference's resumed search
(emp_strategy::set_initial_state + specific_run, include/emp_strategy.h
:332-467), pinned against the uninterrupted search.

Replaying the first k pieces of a situation and restarting at index k - 1
continues the sequential DFS right after the node of the first k - 1 pieces
(index k - 1 is recomputed from its first candidate).  So, with C the number
of nodes up to and including that node in the full search:
  - the resumed node count is N_full - C;
  - every solution found in the full search after node C is found by the
    resumed search, at node (full index - C), and no other;
  - the resumed solutions below the k - 1 pieces are the subtree's
    (oracle_enumerate_subtree, another code path).
C is not computed from the resume: it is read off the two runs and must be
the same for every solution, and at least k - 1.
"""
import numpy as np
import pytest

import edge_matching_puzzle_amd as emp
import oracle_bindings as o


def _situation(w, h, board, k, order="spiral"):
    """The first k strategy positions of board (row-major), others empty."""
    xy = emp.order(w, h, order)
    s = np.full(w * h, 0xFFFF, dtype=np.uint16)
    for x, y in xy[:k]:
        s[y * w + x] = board[y * w + x]
    return s


def _check_suffix(w, h, c, full, res, k):
    full_idx = {tuple(b): int(n) for b, n in zip(full["boards"], full["solution_nodes"])}
    C = full["n_nodes"] - res["n_nodes"]
    assert C >= k - 1
    got = {tuple(b): int(n) for b, n in zip(res["boards"], res["solution_nodes"])}
    want = {b: n - C for b, n in full_idx.items() if n > C}
    assert got == want
    return C


@pytest.mark.parametrize("name", ["3_3", "4_4", "5_5"])
def test_one_piece_is_the_whole_search(example_puzzles, name):
    w, h, c = example_puzzles[name]
    full = o.enumerate_(w, h, c)
    res = o.resume(w, h, c, _situation(w, h, full["boards"][0], 1))
    assert (res["n_nodes"], res["n_solutions"]) == (full["n_nodes"], full["n_solutions"])
    assert list(res["solution_nodes"]) == list(full["solution_nodes"])


@pytest.mark.parametrize("name", ["3_3", "4_4", "5_5"])
def test_resumed_search_is_the_full_search_after_the_prefix(example_puzzles, name):
    w, h, c = example_puzzles[name]
    full = o.enumerate_(w, h, c)
    P = w * h
    for sol in range(len(full["boards"])):
        for k in sorted({2, 3, P // 2, P - 1, P}):
            res = o.resume(w, h, c, _situation(w, h, full["boards"][sol], k))
            _check_suffix(w, h, c, full, res, k)


def _transitions(w, h, c, board, k):
    """Oracle transitions (4 * kind index + o) of the first k positions."""
    xy = emp.order(w, h)
    zeros = (np.asarray(c) == 0).sum(axis=1)
    kinds = [[p for p in range(w * h) if zeros[p] == kk] for kk in range(3)]
    out = []
    for x, y in xy[:k]:
        v = int(board[y * w + x])
        kk = int(zeros[v >> 2])
        out.append(4 * kinds[kk].index(v >> 2) + (v & 3))
    return np.array(out, dtype=np.uint16)


@pytest.mark.parametrize("seed", range(4))
def test_synthetic_resumes_and_their_subtrees(seed):
    """C0 puzzles with identical pieces: the resumed solutions that extend
    the k - 1 pieces are the subtree's solutions."""
    w = h = 5
    c, planted = emp.generate(400 + seed, w, h, 3, 4)
    full = o.enumerate_(w, h, c)
    boards = list(full["boards"]) or [planted]
    for board in boards[:3]:
        for k in (2, 5, 9, 13, 20):
            res = o.resume(w, h, c, _situation(w, h, board, k))
            _check_suffix(w, h, c, full, res, k)
            pre = _transitions(w, h, c, board, k - 1)
            if k - 1 >= 1:
                sub = o.subtree(w, h, c, pre)
                xy = emp.order(w, h)
                cells = [y * w + x for x, y in xy[:k - 1]]
                mine = {tuple(b) for b in res["boards"] if all(b[q] == board[q] for q in cells)}
                assert mine == {tuple(b) for b in sub["boards"]}


def test_refused_situations(example_puzzles):
    w, h, c = example_puzzles["3_3"]
    full = o.enumerate_(w, h, c)
    empty = np.full(9, 0xFFFF, dtype=np.uint16)
    with pytest.raises(ValueError):
        o.resume(w, h, c, empty)
    bad = _situation(w, h, full["boards"][0], 3)
    xy = emp.order(w, h)
    x, y = xy[1]
    bad[y * w + x] ^= 1                  # the second piece turned: not a candidate
    with pytest.raises(ValueError):
        o.resume(w, h, c, bad)
