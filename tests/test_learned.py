"""Learned candidate ordering: features, candidate lists, depth probes and the
ranked enumeration (which must visit exactly the plain enumeration's
situations)."""
import hashlib
import json
import os

import numpy as np
import pytest

import edge_matching_puzzle_amd as emp
from edge_matching_puzzle_amd import learned
import oracle_bindings as o

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _moves_from_transitions(w, h, colours, transitions):
    """Oracle prefixes (kind_index << 2 | o per strategy index) as packed moves."""
    _, by_kind = learned.piece_kinds(colours)
    xy = emp.order(w, h)
    out = []
    for k, t in enumerate(transitions):
        x, y = xy[k]
        K = learned.cell_kind(x, y, w, h)
        out.append(emp.pack_move(y * w + x, int(by_kind[K][int(t) >> 2]), int(t) & 3))
    return out


def _board_hash(boards):
    b = np.asarray(boards, dtype=np.uint16).reshape(len(boards), -1)
    return hashlib.sha256(b"".join(sorted(bytes(r.tobytes()) for r in b))).hexdigest()


def test_feature_layout():
    w, h, c = learned.puzzle(11, 5, 5, 3, 4)
    X = learned.features(c, 25, 3, [(0, 0), (7, 1)], learned.histogram(c, np.ones(25, bool)))
    assert X.shape == (2, emp.MLP_INPUTS)
    for r, (p, o_) in enumerate([(0, 0), (7, 1)]):
        cols = learned.rotated(c, p, o_)
        assert X[r, :88].sum() == sum(1 for v in cols if v >= 1)
        for s, v in enumerate(cols):
            if v >= 1:
                assert X[r, s * 22 + v - 1] == 1
    hist = X[0, 88:110]
    assert abs(hist.sum() - np.count_nonzero(c) / c.size) < 1e-6   # frame edges are not counted
    assert np.allclose(X[0, 110:], learned.position_embedding(3, 25))


def test_candidates_match_oracle_prefixes():
    # the candidates after each depth-d prefix are exactly the oracle's
    # depth-(d+1) prefixes that extend it
    for seed in (3, 8):
        w, h = 5, 5
        c, _ = emp.generate(seed, w, h, 3, 4)
        xy = emp.order(w, h)
        pre3, _, _ = o.prefixes(w, h, c, 3)
        pre4, _, _ = o.prefixes(w, h, c, 4)
        ext = {}
        for row in pre4:
            ext.setdefault(tuple(int(v) for v in row[:3]), []).append(int(row[3]))
        _, by_kind = learned.piece_kinds(c)
        for row in pre3[:40]:
            moves = _moves_from_transitions(w, h, c, row)
            got = learned.candidates_at(w, h, c, xy, moves)
            x, y = xy[3]
            K = learned.cell_kind(x, y, w, h)
            kidx = {int(p): i for i, p in enumerate(by_kind[K])}
            got_t = sorted(4 * kidx[p] + o_ for p, o_ in got)
            assert got_t == sorted(ext.get(tuple(int(v) for v in row), []))


def test_subtree_depth_invariants():
    # depth below a prefix = max over its children (or the prefix itself), and
    # it reaches P exactly when the subtree holds a solution
    w, h = 5, 5
    c, _ = emp.generate(21, w, h, 3, 4)
    pre5, _, _ = o.prefixes(w, h, c, 5)
    pre6, _, _ = o.prefixes(w, h, c, 6)
    kids = {}
    for row in pre6:
        kids.setdefault(tuple(int(v) for v in row[:5]), []).append(row)
    for row in pre5[:30]:
        nodes, depth = o.subtree_depth(w, h, c, row)
        ch = kids.get(tuple(int(v) for v in row), [])
        expect = max([5] + [o.subtree_depth(w, h, c, k)[1] for k in ch])
        assert depth == expect
        assert nodes == len(ch) + sum(o.subtree_depth(w, h, c, k)[0] for k in ch)
        assert (depth == w * h) == (o.subtree(w, h, c, row)["n_solutions"] > 0)


def test_random_rank_tables_are_permutations():
    w, h, c = learned.puzzle(5, 6, 6, 3, 5)
    r = learned.random_rank_table(w, h, c, 1)
    for row in r:
        assert sorted(row.tolist()) == list(range(r.shape[1]))


@pytest.mark.gpu
def test_probe_matches_oracle_subtrees(gpu_context):
    for seed, depth in ((3, 5), (14, 7)):
        w, h = 5, 5
        c, _ = emp.generate(seed, w, h, 3, 4)
        pre, _, _ = o.prefixes(w, h, c, depth)
        rows = np.array([_moves_from_transitions(w, h, c, p) for p in pre[:200]], dtype=np.uint32)
        res, _ = gpu_context.probe(w, h, c, rows, node_limit=0)
        for p, r in zip(pre[:200], res):
            nodes, d = o.subtree_depth(w, h, c, p)
            assert (r["n_nodes"], r["max_depth"], r["complete"]) == (nodes, d, True)


@pytest.mark.gpu
def test_probe_rejects_a_non_candidate(gpu_context):
    w, h = 5, 5
    c, _ = emp.generate(3, w, h, 3, 4)
    pre, _, _ = o.prefixes(w, h, c, 4)
    row = _moves_from_transitions(w, h, c, pre[0])
    row[3] = (row[3] & ~0x30000) | (((row[3] >> 16) + 1) & 3) << 16   # another orientation
    with pytest.raises(emp.EmpError):
        gpu_context.probe(w, h, c, np.array([row], dtype=np.uint32), node_limit=0)


@pytest.mark.gpu
def test_ranked_enumeration_matches_plain(gpu_context):
    cases = [learned.puzzle(s, 5, 5, 3, 4) for s in range(6)] + [learned.puzzle(s, 6, 6, 3, 5) for s in range(3)]
    plain, _ = gpu_context.enumerate(cases, max_solutions=4096)
    ranks = [learned.random_rank_table(w, h, c, 100 + i) for i, (w, h, c) in enumerate(cases)]
    ranked, _ = gpu_context.enumerate(cases, max_solutions=4096, ranks=ranks)
    for (w, h, c), a, b in zip(cases, plain, ranked):
        ref = o.enumerate_(w, h, c, max_solutions=4096)
        assert a["n_nodes"] == b["n_nodes"] == ref["n_nodes"]
        assert a["n_solutions"] == b["n_solutions"] == ref["n_solutions"]
        assert _board_hash(a["boards"]) == _board_hash(b["boards"]) == _board_hash(ref["boards"])
        assert a["max_depth"] == b["max_depth"]


@pytest.mark.gpu
def test_learned_order_exact_on_256_8x8(gpu_context):
    """The 256 8x8 instances enumerated in one batch with a network's visit
    order give the oracle's node counts and solution sets."""
    gold = json.load(open(os.path.join(GOLDEN, "c4_exact_8x8.json")))
    cases = [learned.puzzle(r["seed"], **learned.EXACT_SHAPE) for r in gold["instances"]]
    params = learned.init_params(7)
    ranks = [learned.rank_table(gpu_context, params, w, h, c) for (w, h, c) in cases]
    res, _ = gpu_context.enumerate(cases, max_solutions=4096, ranks=ranks)
    for r, g in zip(res, gold["instances"]):
        assert r["complete"]
        assert (r["n_nodes"], r["n_solutions"]) == (g["n_nodes"], g["n_solutions"]), g["seed"]
        assert (_board_hash(r["boards"]) if r["n_solutions"] else "") == g["boards_sha256"], g["seed"]


def test_row_features_match_per_candidate_features():
    w, h, c = learned.puzzle(learned.TRAIN_SEEDS[0], **learned.C3_SHAPE)
    pre = emp.spiral_prefixes(w, h, c, 7, 6, 3)
    rows = np.array([learned.canonical_moves(c, r) for r in pre], dtype=np.uint32)
    X = learned.row_features(c, w, h, rows)
    for i, r in enumerate(rows):
        remaining = np.ones(w * h, bool)
        for mv in r[:-1]:
            remaining[(int(mv) >> 8) & 0xFF] = False
        p, o_ = (int(r[-1]) >> 8) & 0xFF, (int(r[-1]) >> 16) & 3
        ref = learned.features(c, w * h, len(r) - 1, [(p, o_)], learned.histogram(c, remaining))[0]
        assert np.allclose(X[i], ref)


def test_canonical_moves_are_enumeration_candidates():
    w, h, c = learned.puzzle(learned.TRAIN_SEEDS[1], **learned.C3_SHAPE)
    xy = emp.order(w, h)
    for s in range(6):
        pre = emp.spiral_prefixes(w, h, c, 30, 1, s)
        if len(pre) == 0:
            continue
        mv = learned.canonical_moves(c, pre[0])
        for i in range(len(mv)):
            assert ((mv[i] >> 8) & 0xFF, (mv[i] >> 16) & 3) in learned.candidates_at(w, h, c, xy, mv[:i])
