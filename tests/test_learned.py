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


def _colour_grids_hash(colours, boards, ids=None):
    """Hash of the set of colour layouts of the boards (piece ids mapped back
    through `ids`): equal for two solution sets that differ only by which of
    several identical pieces sits where."""
    c = np.asarray(colours).reshape(-1, 4)
    b = np.asarray(boards, dtype=np.int64).reshape(len(boards), -1)
    p, o_ = b >> 2, b & 3
    if ids is not None:
        p = np.asarray(ids)[p]
    side = np.arange(4)[None, None, :]
    grid = c[p[:, :, None], (side + o_[:, :, None]) & 3].astype(np.uint8)
    return hashlib.sha256(b"".join(sorted(g.tobytes() for g in grid))).hexdigest()


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


def test_expected_histogram_ends_and_sampling():
    """At index 0 every piece remains; past the frame only interior pieces;
    in between, the mean over random kind-respecting placements."""
    w, h, c = learned.puzzle(9, 6, 6, 3, 5)
    P = w * h
    kind, by_kind = learned.piece_kinds(c)
    assert np.allclose(learned.expected_histogram(c, w, h, 0), learned.histogram(c, np.ones(P, bool)))
    frame = 2 * (w + h) - 4
    inner = kind == 0
    for k in (frame, frame + 3, P - 1):
        assert np.allclose(learned.expected_histogram(c, w, h, k), learned.histogram(c, inner), atol=1e-6)
    xy = emp.order(w, h)
    rng = np.random.default_rng(3)
    k = 9
    acc = np.zeros(learned.N_COLOURS)
    trials = 3000
    for _ in range(trials):
        remaining = np.ones(P, bool)
        pools = {K: list(rng.permutation(by_kind[K])) for K in range(3)}
        for x, y in xy[:k]:
            remaining[pools[learned.cell_kind(x, y, w, h)].pop()] = False
        acc += learned.histogram(c, remaining)
    assert np.abs(acc / trials - learned.expected_histogram(c, w, h, k)).max() < 4e-3


def test_renumbering_keeps_the_tree():
    """Renumbered pieces with the first corner kept give the oracle's node
    count and solution layouts of the original numbering; moving another
    corner to the front searches another tree."""
    for seed in range(4):
        w, h, c = learned.puzzle(seed, 5, 5, 3, 4)
        ref = o.enumerate_(w, h, c, max_solutions=4096)
        rc, ids = learned.shuffled(c, 10 + seed)
        assert sorted(ids.tolist()) == list(range(w * h))
        assert np.array_equal(np.asarray(c).reshape(-1, 4)[ids], rc)
        got = o.enumerate_(w, h, rc, max_solutions=4096)
        assert (got["n_nodes"], got["n_solutions"]) == (ref["n_nodes"], ref["n_solutions"])
        assert _colour_grids_hash(c, got["boards"], ids) == _colour_grids_hash(c, ref["boards"])
    w, h, c = learned.puzzle(0, 5, 5, 3, 4)
    kind, by_kind = learned.piece_kinds(c)
    prio = np.zeros(w * h)
    prio[by_kind[2][1]] = 1.0
    perm = np.lexsort((np.arange(w * h), -prio))    # second corner first
    other = o.enumerate_(w, h, np.asarray(c).reshape(-1, 4)[perm])
    assert other["n_nodes"] != o.enumerate_(w, h, c)["n_nodes"]


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


TRAIN_SEEDS = learned.TRAIN_SEEDS[:8]


@pytest.fixture(scope="module")
def trained(gpu_context):
    """A scorer trained on depth-probe traces of eight 16x16 training seeds
    (64 situations each, 5,000-placement probes), with its held-out error."""
    print("case: training the scorer on %d seeds" % len(TRAIN_SEEDS), flush=True)
    Xs, ys = [], []
    for seed in TRAIN_SEEDS:
        w, h, c = learned.puzzle(seed, **learned.C3_SHAPE)
        X, y, _ = learned.traces(gpu_context, w, h, c, 64, seed, 5000, depths=(4, 180))
        Xs.append(X)
        ys.append(y)
    X, y = np.concatenate(Xs), np.concatenate(ys)
    init = learned.init_params(1)
    params, losses = gpu_context.mlp_train(X, y, init, epochs=20, batch=256, lr=1e-3, seed=1)
    hx, hy = [], []
    for seed in learned.HELDOUT_SEEDS[:2]:
        w, h, c = learned.puzzle(seed, **learned.C3_SHAPE)
        Xh, yh, _ = learned.traces(gpu_context, w, h, c, 16, seed + 7, 5000, depths=(4, 180))
        hx.append(Xh)
        hy.append(yh)
    Xh, yh = np.concatenate(hx), np.concatenate(hy)
    mse = float(np.mean((gpu_context.mlp_score(params, Xh) - yh) ** 2))
    mse_mean = float(np.mean((y.mean() - yh) ** 2))
    return {"params": params, "init": init, "losses": losses, "rows": len(y), "heldout_mse": mse,
            "heldout_mse_of_mean": mse_mean}


@pytest.mark.gpu
def test_scorer_is_trained(trained):
    """The scorer the ordering tests use has learned: its training loss fell
    and on held-out traces it predicts the depth gained better than the
    training mean does."""
    assert trained["rows"] > 400
    assert not np.array_equal(trained["params"], trained["init"])
    assert trained["losses"][-1] < 0.5 * trained["losses"][0]
    assert trained["heldout_mse"] < trained["heldout_mse_of_mean"], trained


@pytest.mark.gpu
def test_learned_order_exact_on_256_8x8(gpu_context, trained):
    """The 256 8x8 instances enumerated in one batch with the trained
    network's visit order give the oracle's node counts and solution sets;
    renumbered by the trained scores, the same counts and solution layouts."""
    gold = json.load(open(os.path.join(GOLDEN, "c4_exact_8x8.json")))
    cases = [learned.puzzle(r["seed"], **learned.EXACT_SHAPE) for r in gold["instances"]]
    params = trained["params"]
    ranks = [learned.rank_table(gpu_context, params, w, h, c) for (w, h, c) in cases]
    plain_order = sum(int(np.array_equal(r, learned.rank_table(gpu_context, trained["init"], w, h, c)))
                      for r, (w, h, c) in zip(ranks[:8], cases[:8]))
    assert plain_order == 0   # the trained order is not the untrained one
    res, _ = gpu_context.enumerate(cases, max_solutions=4096, ranks=ranks)
    for r, g in zip(res, gold["instances"]):
        assert r["complete"]
        assert (r["n_nodes"], r["n_solutions"]) == (g["n_nodes"], g["n_solutions"]), g["seed"]
        assert (_board_hash(r["boards"]) if r["n_solutions"] else "") == g["boards_sha256"], g["seed"]
    relab = [learned.relabelled(gpu_context, params, w, h, c) for (w, h, c) in cases]
    res2, _ = gpu_context.enumerate([(w, h, rc) for (w, h, _), (rc, _) in zip(cases, relab)], max_solutions=4096)
    for (w, h, c), (rc, ids), r, r2, g in zip(cases, relab, res, res2, gold["instances"]):
        assert r2["complete"]
        assert (r2["n_nodes"], r2["n_solutions"]) == (g["n_nodes"], g["n_solutions"]), g["seed"]
        if r["n_solutions"]:
            assert _colour_grids_hash(c, r2["boards"], ids) == _colour_grids_hash(c, r["boards"]), g["seed"]


@pytest.mark.gpu
def test_trained_order_depth_on_heldout(gpu_context, trained):
    """Max depth reached in equal device time (300 ms) on four held-out
    16x16 seeds by the plain order, the trained order (ranked kernel and
    pieces renumbered by score) and random orders.  The depths are reported
    whatever they are; the search itself is checked: no run completes (these
    instances take far more than 300 ms), each reaches past the frame and none
    places more pieces than the board holds."""
    params = trained["params"]
    report = []
    for seed in learned.HELDOUT_SEEDS[:4]:
        w, h, c = learned.puzzle(seed, **learned.C3_SHAPE)
        P = w * h
        relab, _ = learned.relabelled(gpu_context, params, w, h, c)
        shuffled, _ = learned.shuffled(c, seed)
        row = {"seed": seed}
        for name, cc, rk in (("plain", c, None), ("learned", c, [learned.rank_table(gpu_context, params, w, h, c)]),
                             ("learned_relabel", relab, None),
                             ("random", c, [learned.random_rank_table(w, h, c, seed)]),
                             ("random_relabel", shuffled, None)):
            print("case: held-out seed", seed, name, flush=True)
            (r,), st = gpu_context.enumerate([(w, h, cc)], max_solutions=0, ranks=rk, time_limit_ms=300.0)
            assert not r["complete"] and 2 * (w + h) - 4 < r["max_depth"] <= P, (seed, name, r)
            row[name] = r["max_depth"]
        report.append(row)
    means = {k: float(np.mean([r[k] for r in report])) for k in report[0] if k != "seed"}
    print("held-out mean max depth:", json.dumps(means), "per seed:", json.dumps(report), flush=True)


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
