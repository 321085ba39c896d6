"""The board-validity checker (tests/board_check.py), pinned on boards whose
validity is known by construction, then applied to the CPU oracle's glutton
situations.

The planted solution of a generated puzzle is valid by construction; each
planted mistake (a piece turned, two pieces swapped, a piece used twice, a
wrong depth or solved flag, a prefix move changed) must be caught by both the
scalar and the vectorised form.
"""
import numpy as np
import pytest

import edge_matching_puzzle_amd as emp
import oracle_bindings as o
from board_check import EMPTY, board_problems, invalid_boards


def _shows_differently(c, v, turn):
    q, r = v >> 2, v & 3
    return any(c[q][(s + r) % 4] != c[q][(s + r + turn) % 4] for s in range(4))


@pytest.mark.parametrize("w,h,B,I,seed", [(5, 5, 3, 4, 1), (8, 8, 4, 8, 2), (7, 4, 2, 3, 3), (16, 16, 5, 17, 3)])
def test_planted_solutions_are_valid(w, h, B, I, seed):
    c, planted = emp.generate(seed, w, h, B, I)
    assert board_problems(w, h, c, planted, depth=w * h, solved=True) == []
    assert len(invalid_boards(w, h, c, planted[None, :], depths=[w * h], solved=[True])) == 0


def _turn(b, i, k):
    v = int(b[i])
    b[i] = (v & ~3) | ((v + k) & 3)


def _mutations(w, h, c, planted, rng):
    """(name, board, kwargs) of boards that each break one rule."""
    P = w * h
    out = []
    interior = [i for i in range(P) if 0 < i % w < w - 1 and 0 < i // w < h - 1]
    cell = next(i for i in interior if _shows_differently(c, int(planted[i]), 1))
    b = planted.copy()
    _turn(b, cell, 1)
    out.append(("turned interior piece", b, {}))
    corner = 0
    b = planted.copy()
    _turn(b, corner, 2)
    out.append(("corner facing inwards", b, {}))
    a, z = interior[0], interior[-1]
    b = planted.copy()
    b[a], b[z] = b[z], b[a]
    if board_problems(w, h, c, b):
        out.append(("two pieces swapped", b, {}))
    b = planted.copy()
    b[z] = EMPTY
    b[a] = EMPTY
    b[z] = planted[interior[1]]          # a piece in two cells, both isolated from a's hole
    out.append(("piece used twice", b, {}))
    b = planted.copy()
    b[cell] = EMPTY
    out.append(("wrong depth", b, {"depth": P}))
    out.append(("full board not solved", planted.copy(), {"solved": False}))
    b = planted.copy()
    b[cell] = EMPTY
    out.append(("partial board solved", b, {"solved": True}))
    v = int(planted[interior[2]])
    out.append(("prefix move missing", planted.copy(),
                {"prefix_moves": [emp.pack_move(interior[3], v >> 2, v & 3)]}))
    return out


@pytest.mark.parametrize("w,h,B,I,seed", [(5, 5, 3, 4, 4), (6, 6, 3, 6, 5), (9, 5, 3, 5, 6)])
def test_planted_mistakes_are_caught(w, h, B, I, seed):
    c, planted = emp.generate(seed, w, h, B, I)
    rng = np.random.default_rng(seed)
    cases = _mutations(w, h, c, planted, rng)
    assert len(cases) >= 7
    for name, b, kw in cases:
        assert board_problems(w, h, c, b, **kw), name
        vkw = {}
        if "depth" in kw:
            vkw["depths"] = [kw["depth"]]
        if "solved" in kw:
            vkw["solved"] = [kw["solved"]]
        if "prefix_moves" in kw:
            vkw["prefixes"] = np.array([kw["prefix_moves"]], dtype=np.uint32)
        assert list(invalid_boards(w, h, c, b[None, :], **vkw)) == [0], name


def test_vectorised_form_agrees_on_a_batch():
    w = h = 6
    c, planted = emp.generate(9, w, h, 3, 5)
    boards = []
    rng = np.random.default_rng(9)
    for _ in range(64):
        b = planted.copy()
        k = rng.integers(0, 3)
        if k == 1:
            i = int(rng.integers(0, w * h))
            _turn(b, i, int(rng.integers(1, 4)))
        elif k == 2:
            i, j = rng.integers(0, w * h, size=2)
            b[i], b[j] = b[j], b[i]
        drop = rng.integers(0, w * h, size=int(rng.integers(0, 10)))
        b[drop] = EMPTY
        boards.append(b)
    boards = np.array(boards)
    bad = set(invalid_boards(w, h, c, boards).tolist())
    for i, b in enumerate(boards):
        assert (i in bad) == bool(board_problems(w, h, c, b)), i
    assert 0 < len(bad) < 64


def _triples(row):
    return [v for m in row for v in ((int(m) & 0xFF), (int(m) >> 8) & 0xFF, (int(m) >> 16) & 3)]


@pytest.mark.parametrize("w,h,B,I,seed,length,budget", [(5, 5, 3, 4, 1, 4, 200), (6, 6, 3, 6, 2, 6, 80),
                                                         (7, 5, 2, 3, 3, 5, 150), (8, 8, 4, 8, 4, 10, 30)])
def test_oracle_glutton_situations_are_valid(w, h, B, I, seed, length, budget):
    """The oracle's deepest situations, from the empty board and from
    prefixes: edge-valid, each piece once, prefix kept, depth = pieces."""
    c, _ = emp.generate(seed, w, h, B, I)
    r = o.glutton_max(w, h, c, push_limit=budget)
    assert board_problems(w, h, c, r["board"], depth=r["max_depth"], solved=r["solved"]) == []
    for row in emp.spiral_prefixes(w, h, c, length, 6, seed):
        r = o.glutton_max(w, h, c, prefix=_triples(row), push_limit=budget)
        assert board_problems(w, h, c, r["board"], prefix_moves=row, depth=r["max_depth"],
                              solved=r["solved"]) == []
        # a prefix whose last move leaves a cell without option is not
        # searched: its deepest situation is the moves before that one
        assert r["max_depth"] >= length or (r["pushes"] == 0 and r["status"] == o.OK and not r["solved"])


def test_golden_16x16_situations_are_valid(golden_dir):
    import json
    import os
    g = json.load(open(os.path.join(golden_dir, "glutton_16x16_seed3.json")))
    c, _ = emp.generate(g["seed"], 16, 16, g["border_colours"], g["interior_colours"])
    for s in g["searches"]:
        assert board_problems(16, 16, c, s["board"], prefix_moves=s["prefix"], depth=s["max_depth"],
                              solved=s["solved"]) == []
