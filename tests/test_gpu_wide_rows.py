"""Enumeration with mask rows wider than four words (boards above 10x10: u16
stacks, tables read through the caches), against the oracle: the subtree
below a situation deep inside a planted solution (the search of the
remaining positions with the situation's pieces fixed) must have the
oracle's node count and solution set."""
import numpy as np
import pytest

import edge_matching_puzzle_amd as emp
import oracle_bindings as o

pytestmark = pytest.mark.gpu


def _matches(w, h, colours, board):
    c = np.asarray(colours).reshape(-1, 4)
    shown = lambda v, side: int(c[v >> 2][(side + (v & 3)) % 4])
    for y in range(h):
        for x in range(w):
            v = int(board[y * w + x])
            if x + 1 < w and shown(v, 1) != shown(int(board[y * w + x + 1]), 3):
                return False
            if y + 1 < h and shown(v, 2) != shown(int(board[(y + 1) * w + x]), 0):
                return False
    return True


def _in_the_tree(w, h, colours, board):
    """A square solution turned by quarter turns until the corner the
    enumeration fixes at (0, 0) -- the first corner piece -- is there (the
    search visits only that corner at index 0)."""
    c = np.asarray(colours).reshape(-1, 4)
    first_corner = int(np.flatnonzero((c == 0).sum(axis=1) == 2)[0])
    b = np.asarray(board, dtype=np.int64)
    for _ in range(4):
        if int(b[0]) >> 2 == first_corner:
            return b.astype(np.uint16)
        turned = None
        for d in (1, 3):   # the orientation step of a quarter turn, whichever matches
            t = np.empty_like(b)
            for y in range(h):
                for x in range(w):
                    v = int(b[y * w + x])
                    t[x * w + (w - 1 - y)] = (v & ~3) | (((v & 3) + d) & 3)
            if _matches(w, h, colours, t):
                turned = t
                break
        assert turned is not None
        b = turned
    raise AssertionError("no turn puts the first corner at (0, 0)")


def _canonical(w, h, colours, board):
    """The same solution with identical pieces (the same four colours up to
    rotation) exchanged so that, in spiral order, each placement uses the
    lowest-numbered one not placed yet -- the only one the enumeration tries
    -- turned to show the same colours, and pieces with four equal sides at
    orientation 0."""
    c = np.asarray(colours).reshape(-1, 4)
    P = len(c)
    look = lambda q, o: tuple(int(c[q][(side + o) % 4]) for side in range(4))
    key = [min(look(q, o) for o in range(4)) for q in range(P)]
    b = np.asarray(board, dtype=np.int64).copy()
    cell_of = {int(b[i]) >> 2: i for i in range(P)}
    placed = set()
    for x, y in emp.order(w, h):
        cell = y * w + x
        p, o = int(b[cell]) >> 2, int(b[cell]) & 3
        low = min(q for q in range(P) if key[q] == key[p] and q not in placed)
        if low != p:
            other = cell_of[low]
            o_low = int(b[other]) & 3
            b[cell] = low << 2 | next(t for t in range(4) if look(low, t) == look(p, o))
            b[other] = p << 2 | next(t for t in range(4) if look(p, t) == look(low, o_low))
            cell_of[low], cell_of[p] = cell, other
            p = low
        if len(set(c[p])) == 1:
            b[cell] = p << 2
        placed.add(p)
    return b.astype(np.uint16)


def _prefix(w, h, colours, board, k):
    """The first k spiral placements of a board as moves (pos | piece << 8 |
    o << 16, padded to P) and as the oracle's transitions (kind index << 2 |
    o, the kind index being the piece's rank among the pieces of its kind)."""
    colours = np.asarray(colours).reshape(-1, 4)
    kind = (colours == 0).sum(axis=1)
    rank = np.zeros(len(colours), dtype=np.int64)
    for kd in range(3):
        ids = np.flatnonzero(kind == kd)
        rank[ids] = np.arange(len(ids))
    moves = np.zeros(w * h, dtype=np.uint32)
    trans = np.zeros(k, dtype=np.uint16)
    for i, (x, y) in enumerate(emp.order(w, h)[:k]):
        v = int(board[y * w + x])
        piece, ori = v >> 2, v & 3
        moves[i] = emp.pack_move(y * w + x, piece, ori)
        trans[i] = (rank[piece] << 2) | ori
    return moves, trans


# free positions chosen so that each subtree holds 10^6 to 3*10^7 nodes (the
# oracle enumerates each in under two seconds); row widths 6 (11x11), 8
# (12x12, 13x13) and 16 (16x16) words
@pytest.mark.parametrize("w,h,B,I,seed,free,schedule", [
    (11, 11, 4, 9, 5, 60, "auto"), (12, 12, 5, 10, 6, 64, "auto"), (12, 12, 5, 10, 8, 64, "passes"),
    (13, 13, 5, 12, 8, 80, "auto"), (16, 16, 5, 17, 3, 128, "auto"), (16, 16, 5, 17, 3, 128, "passes"),
    (16, 16, 5, 17, 5, 124, "persistent")])
def test_subtree_below_a_planted_situation_matches_the_oracle(gpu_context, w, h, B, I, seed, free, schedule):
    c, planted = emp.generate(seed, w, h, B, I)
    P = w * h
    k = P - free
    board = _canonical(w, h, c, _in_the_tree(w, h, c, planted))
    moves, trans = _prefix(w, h, c, board, k)
    ref = o.subtree(w, h, c, trans, max_solutions=1 << 12)
    print("case: wide rows", w, h, B, I, seed, k, schedule, ref["n_nodes"], flush=True)
    r = gpu_context.enumerate_items(w, h, c, items=[[k, k, emp.NO_MOVE]], moves=[moves], max_solutions=1 << 12,
                                    schedule=schedule)
    assert r["stats"]["max_words"] > 4
    assert r["complete"]
    assert (r["n_nodes"], r["n_solutions"]) == (ref["n_nodes"], ref["n_solutions"])
    assert ref["n_solutions"] >= 1   # the planted solution is below the situation
    assert sorted(map(tuple, r["boards"].tolist())) == sorted(map(tuple, ref["boards"].tolist()))
