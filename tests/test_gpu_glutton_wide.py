"""glutton_wide's breadth step (emp_hip_glutton_wide, through the C ABI).

Bars: the level-0 masks, each child's masks and each status equal the
oracle's (moves replayed by its glutton restatement, whose move application
the glutton known answers pin); the children are the parent cell's mask bits
in order; and, independently of the oracle, every option left at a free
position agrees with the colours its placed neighbours show it, uses no
placed piece and faces the frame with colour 0.
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


def _triples(moves):
    return [v for m in moves for v in (int(m) & 0xFF, (int(m) >> 8) & 0xFF, (int(m) >> 16) & 3)]


def _bits(mask):
    out = []
    for w, v in enumerate(mask):
        v = int(v)
        while v:
            b = (v & -v).bit_length() - 1
            out.append(64 * w + b)
            v &= v - 1
    return out


def _options_consistent(w, h, c, moves, masks):
    """Every (piece, o) left at a free position fits its placed neighbours
    and the frame, and no placed piece is left anywhere."""
    P = w * h
    board = {int(m) & 0xFF: ((int(m) >> 8) & 0xFF, (int(m) >> 16) & 3) for m in moves}
    used = {p for p, _ in board.values()}
    for q in range(P):
        bits = _bits(masks[q])
        if q in board:
            assert not bits
            continue
        x, y = q % w, q // w
        for b in bits:
            piece, r = b // 4, b % 4
            assert piece not in used
            for s, (dx, dy) in enumerate(((0, -1), (1, 0), (0, 1), (-1, 0))):
                col = int(c[piece][(s + r) % 4])
                nx, ny = x + dx, y + dy
                if not (0 <= nx < w and 0 <= ny < h):
                    assert col == 0
                    continue
                nb = ny * w + nx
                if nb in board:
                    npiece, nr = board[nb]
                    assert col == int(c[npiece][((s + 2) % 4 + nr) % 4])


def _check(ctx, w, h, c, parents, cells):
    res = ctx.glutton_wide(w, h, c, parents=parents, cells=cells)
    n = len(cells)
    rows = parents if parents is not None else np.zeros((1, 0), dtype=np.uint32)
    expected = 0
    for p in range(n):
        ref, st = o.glutton_masks(w, h, c, _triples(rows[p]))
        assert res["parent_status"][p] == st
        if st == 0:
            assert (res["parent_masks"][p] == ref).all()
            expected += len(_bits(ref[cells[p]]))
    assert res["n_children"] == expected
    for i in range(res["n_children"]):
        p, mv = int(res["child_parent"][i]), int(res["child_move"][i])
        assert mv & 0xFF == cells[p]
        moves = list(rows[p]) + [mv]
        ref, st = o.glutton_masks(w, h, c, _triples(moves))
        assert res["child_status"][i] == st
        if st == 0:
            assert (res["child_masks"][i] == ref).all()
            if w * h <= 64:   # a Python loop over every option
                _options_consistent(w, h, c, moves, res["child_masks"][i])
    # children of a parent come in bit order
    for p in range(n):
        got = [4 * ((int(m) >> 8) & 0xFF) + ((int(m) >> 16) & 3)
               for q, m in zip(res["child_parent"], res["child_move"]) if q == p]
        assert got == sorted(got)
    return res


@pytest.mark.parametrize("name", ["3_3", "4_4", "5_5"])
def test_corner_expansion_of_the_examples(gpu_context, example_puzzles, name):
    """The reference's run: the level-0 situation expanded at (0, 0)."""
    w, h, c = example_puzzles[name]
    print("case: glutton_wide corner", name, flush=True)
    res = _check(gpu_context, w, h, c, None, [0])
    assert res["n_children"] == 4          # four corners, one orientation each


@pytest.mark.parametrize("w,h,B,I,seed,length", [(5, 5, 3, 4, 1, 4), (6, 6, 3, 6, 2, 9), (8, 8, 4, 8, 3, 12),
                                                 (7, 5, 2, 3, 4, 6), (16, 16, 5, 17, 3, 24)])
def test_prefixes_expanded_at_the_next_cell(gpu_context, w, h, B, I, seed, length):
    c, _ = emp.generate(seed, w, h, B, I)
    rows = emp.spiral_prefixes(w, h, c, length, 6, seed)
    order = emp.order(w, h)
    nx, ny = order[length]
    cells = [ny * w + nx] * len(rows)
    print("case: glutton_wide prefixes", w, h, length, len(rows), flush=True)
    _check(gpu_context, w, h, c, rows, cells)


def test_unplayable_parent_move(gpu_context):
    w = h = 5
    c, _ = emp.generate(3, w, h, 3, 4)
    rows = emp.spiral_prefixes(w, h, c, 4, 2, 3).copy()
    rows[1, 2] = rows[1, 1]                # the same cell twice
    res = gpu_context.glutton_wide(w, h, c, parents=rows, cells=[20, 20])
    assert list(res["parent_status"]) == [emp.WIDE_VALID, emp.WIDE_BAD_MOVE]
    assert all(p == 0 for p in res["child_parent"])


def test_glutton_wide_command_prints_the_corner_children(example_puzzles):
    """The command's output has the reference's shape: the level-0
    situation, the corner info, then each candidate bit and its child."""
    w, h, c = example_puzzles["3_3"]
    print("case: CLI CUDA_glutton_wide", flush=True)
    p = subprocess.run([EXE, os.path.join(PUZZLES, "3_3.txt"), "ressources", "CUDA_glutton_wide",
                        "--enable_gui=false"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout
    lines = p.stdout.splitlines()
    assert lines.count("Situation Level 0: ") == 1 and lines.count("Situation Level 1: ") == 4
    assert "Corner info : " in lines
    level0, _ = o.glutton_masks(w, h, c, [])
    bits = _bits(level0[0])
    start = lines.index("Corner info : ")
    numbers = [int(l) for l in lines[start:] if l.strip().isdigit()]
    assert numbers == [b + 1 for b in bits]
    assert "Position[0] -> Index 4294967295" in lines
    for i in range(4):
        assert "Situation Index %d: " % i in lines
