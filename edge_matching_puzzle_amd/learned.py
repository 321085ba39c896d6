"""Learned candidate ordering.

The search visits the candidates of a position in an order; the plain order
is the reference's (kind index, then orientation: include/emp_strategy.h
:332-415).  A small network scores each (situation, candidate) and the
ranked enumeration (emp_hip_enumerate_ranked) visits the candidates of each
strategy index by decreasing score.  Ordering never removes a candidate, so
the situations visited, the node count and the solution set of a complete
enumeration are unchanged; only how soon the deep situations are reached
changes.

Features of a candidate (piece p at orientation o, at strategy index k):
  * its four edge colours as placed (north, east, south, west), one-hot over
    colours 1..22 (the frame colour 0 sets no bit): 88 inputs;
  * the colour histogram of the pieces still to place (edges of colour c over
    all their edges): 22 inputs.  Traces use the situation's own remaining
    pieces; a static rank table, which serves every situation at index k,
    uses their expectation at k (expected_histogram);
  * the strategy index, sin / cos of pi 2^i k / P for i = 0..7: 16 inputs.
Labels: the most pieces placed in one situation of the plain search below
the candidate within a node budget (emp_hip_probe), less the pieces placed
with it, over 64.
"""
import numpy as np

import edge_matching_puzzle_amd as emp

N_COLOURS = 22
N_FEATURES = 4 * N_COLOURS + N_COLOURS + 16
assert N_FEATURES == emp.MLP_INPUTS

# Shapes of the learned-ordering workload.
C3_SHAPE = dict(width=16, height=16, border_colours=5, interior_colours=17)
EXACT_SHAPE = dict(width=8, height=8, border_colours=8, interior_colours=10)
TRAIN_SEEDS = [4_000_000 + i for i in range(64)]
HELDOUT_SEEDS = [4_100_000 + i for i in range(16)]
EXACT_SEEDS = [5_000_000 + i for i in range(256)]


def puzzle(seed, width, height, border_colours, interior_colours):
    colours, _ = emp.generate(seed, width, height, border_colours, interior_colours)
    return width, height, colours


def piece_kinds(colours):
    """Kind of each piece (number of frame sides) and the pieces of each kind
    in id order (kind index = rank in that list)."""
    colours = np.asarray(colours).reshape(-1, 4)
    kind = (colours == 0).sum(1)
    return kind, [np.flatnonzero(kind == k) for k in range(3)]


def cell_kind(x, y, width, height):
    return int(x == 0) + int(x == width - 1) + int(y == 0) + int(y == height - 1)


def rotated(colours, piece, o):
    """Colours shown north, east, south, west by `piece` at orientation o."""
    c = np.asarray(colours).reshape(-1, 4)[piece]
    return [int(c[(s + o) & 3]) for s in range(4)]


def orientations_kept(colours, piece):
    """Orientations the reference tries: one for a piece equal under a
    quarter turn, 0 and 2 for one equal under a half turn, else all four
    (include/emp_piece_db.h:1266-1272)."""
    c = np.asarray(colours).reshape(-1, 4)[piece]
    if c[0] == c[1] == c[2] == c[3]:
        return (0,)
    if c[0] == c[2] and c[1] == c[3]:
        return (0, 2)
    return (0, 1, 2, 3)


def histogram(colours, remaining):
    """Edge colour histogram of the pieces in `remaining` (bool mask)."""
    c = np.asarray(colours).reshape(-1, 4)[np.asarray(remaining, dtype=bool)].reshape(-1)
    h = np.zeros(N_COLOURS, dtype=np.float32)
    c = c[(c >= 1) & (c <= N_COLOURS)]
    np.add.at(h, c - 1, 1.0)
    total = max(4 * int(np.count_nonzero(remaining)), 1)
    return h / total


def expected_histogram(colours, width, height, k):
    """Expected colour histogram of the pieces still to place at strategy
    index k: a cell of kind K takes a piece of kind K, so kind K keeps
    (cells of kind K at indices >= k) / (cells of kind K) of its pieces.  On
    the spiral the frame comes first, so past the frame only interior
    colours remain."""
    c = np.asarray(colours).reshape(-1, 4)
    kind, _ = piece_kinds(c)
    order_xy = emp.order(width, height)
    left = np.zeros(3)
    total = np.zeros(3)
    for i, (x, y) in enumerate(order_xy):
        K = cell_kind(x, y, width, height)
        total[K] += 1
        left[K] += i >= k
    h = np.zeros(N_COLOURS, dtype=np.float64)
    edges = 0.0
    for K in range(3):
        if total[K] == 0:
            continue
        share = left[K] / total[K]
        ck = c[kind == K].reshape(-1)
        edges += 4 * share * np.count_nonzero(kind == K)
        ck = ck[(ck >= 1) & (ck <= N_COLOURS)]
        np.add.at(h, ck - 1, share)
    return (h / max(edges, 1.0)).astype(np.float32)


def position_embedding(k, P):
    f = np.pi * (2.0 ** np.arange(8)) * (k / P)
    return np.concatenate([np.sin(f), np.cos(f)]).astype(np.float32)


def features(colours, P, k, candidates, hist):
    """Rows for candidates [(piece, o), ...] at strategy index k."""
    out = np.zeros((len(candidates), N_FEATURES), dtype=np.float32)
    pe = position_embedding(k, P)
    for r, (p, o) in enumerate(candidates):
        for s, col in enumerate(rotated(colours, p, o)):
            if 1 <= col <= N_COLOURS:
                out[r, s * N_COLOURS + col - 1] = 1.0
        out[r, 4 * N_COLOURS:5 * N_COLOURS] = hist
        out[r, 5 * N_COLOURS:] = pe
    return out


def candidates_at(width, height, colours, order_xy, moves):
    """Candidates of the enumeration at strategy index len(moves) after the
    packed moves, as (piece, o): pieces of the cell's kind not placed, kept
    orientations matching the frame and the placed neighbours; of the
    candidates that look alike (same four colours) only the lowest piece's
    are kept (include/emp_position_strategy.h:189-202)."""
    colours = np.asarray(colours).reshape(-1, 4)
    kind, by_kind = piece_kinds(colours)
    k = len(moves)
    x, y = order_xy[k]
    K = cell_kind(x, y, width, height)
    board = {}
    used = set()
    for mv in moves:
        pos, p, o = mv & 0xFF, (mv >> 8) & 0xFF, (mv >> 16) & 3
        board[pos] = (p, o)
        used.add(p)
    need = []
    for s, (dx, dy) in enumerate([(0, -1), (1, 0), (0, 1), (-1, 0)]):
        nx, ny = x + dx, y + dy
        if nx < 0 or ny < 0 or nx >= width or ny >= height:
            need.append(0)
        elif ny * width + nx in board:
            q, qo = board[ny * width + nx]
            need.append(rotated(colours, q, qo)[(s + 2) & 3])
        else:
            need.append(None)
    out, seen = [], {}
    for p in by_kind[K]:
        if p in used:
            continue
        for o in orientations_kept(colours, p):
            look = rotated(colours, p, o)
            ok = all((need[s] is None and look[s] != 0) or (need[s] is not None and look[s] == need[s])
                     for s in range(4))
            if not ok:
                continue
            key = tuple(look)
            if key in seen and seen[key] != p:
                continue
            seen.setdefault(key, p)
            out.append((int(p), o))
    return out


def canonical_moves(colours, moves):
    """The same situation with identical pieces relabelled so that every move
    uses the lowest piece still available among those showing the same
    colours (the enumeration only tries that one,
    include/emp_position_strategy.h:189-202).  Identical pieces are rotations
    of each other, so swapping two of them everywhere leaves the board's
    colours unchanged."""
    colours = np.asarray(colours).reshape(-1, 4)
    kind, by_kind = piece_kinds(colours)
    moves = [int(m) for m in moves]

    def key(p):
        c = [int(v) for v in colours[p]]
        return min(tuple(c[r:] + c[:r]) for r in range(4))

    used = set()
    for i, mv in enumerate(moves):
        pos, p, o = mv & 0xFF, (mv >> 8) & 0xFF, (mv >> 16) & 3
        look = rotated(colours, p, o)
        for q in by_kind[kind[p]]:
            q = int(q)
            if q in used or key(q) != key(p):
                continue
            if q != p:
                # q shows the same look at some orientation; swap p and q in
                # every later move
                oq = next(r for r in orientations_kept(colours, q) if rotated(colours, q, r) == look)
                for j in range(i + 1, len(moves)):
                    pj, oj = (moves[j] >> 8) & 0xFF, (moves[j] >> 16) & 3
                    if pj == q:
                        lj = rotated(colours, q, oj)
                        op = next(r for r in orientations_kept(colours, p) if rotated(colours, p, r) == lj)
                        moves[j] = emp.pack_move(moves[j] & 0xFF, p, op)
                moves[i] = emp.pack_move(pos, q, oq)
                p = q
            break
        used.add(p)
    return moves


def traces(ctx, width, height, colours, n_situations, seed, node_limit, depths=(4, 200)):
    """(features, labels, depth reached) from depth probes of the plain search.

    Situations are random spiral prefixes of random length in `depths`; each
    of their candidates is probed with node_limit placements."""
    P = width * height
    order_xy = emp.order(width, height)
    rng = np.random.default_rng(seed)
    rows, feats, base = [], [], []
    lengths = rng.integers(depths[0], min(depths[1], P - 2), n_situations)
    for i, L in enumerate(lengths):
        pre = emp.spiral_prefixes(width, height, colours, int(L), 1, int(rng.integers(1 << 62)))
        if len(pre) == 0:
            continue
        moves = canonical_moves(colours, pre[0])
        cands = candidates_at(width, height, colours, order_xy, moves)
        if not cands:
            continue
        remaining = np.ones(P, dtype=bool)
        for mv in moves:
            remaining[(mv >> 8) & 0xFF] = False
        x, y = order_xy[L]
        cell = y * width + x
        for p, o in cands:
            rows.append(moves + [emp.pack_move(cell, p, o)])
            base.append(L + 1)
        feats.append(features(colours, P, L, cands, histogram(colours, remaining)))
    if not rows:
        return np.zeros((0, N_FEATURES), np.float32), np.zeros(0, np.float32), np.zeros(0, np.int64)
    X = np.concatenate(feats)
    y = np.zeros(len(rows), dtype=np.float32)
    depth = np.zeros(len(rows), dtype=np.int64)
    # probes grouped by row length (one launch per length)
    by_len = {}
    for r, row in enumerate(rows):
        by_len.setdefault(len(row), []).append(r)
    for L, idx in by_len.items():
        res, _ = ctx.probe(width, height, colours, np.array([rows[r] for r in idx], dtype=np.uint32), node_limit)
        for r, out in zip(idx, res):
            depth[r] = out["max_depth"]
            y[r] = (out["max_depth"] - base[r]) / 64.0
    return X, y, depth


def init_params(seed):
    """He-style initial parameters in the C ABI's layout."""
    rng = np.random.default_rng(seed)
    IN_PAD, HID = 128, 256
    p = np.zeros(emp.MLP_PARAMS, dtype=np.float32)
    w1 = rng.normal(0, np.sqrt(2.0 / N_FEATURES), (IN_PAD, HID)).astype(np.float32)
    w1[N_FEATURES:] = 0
    o = 0
    p[o:o + IN_PAD * HID] = w1.reshape(-1); o += IN_PAD * HID
    p[o:o + HID] = 0.01; o += HID
    p[o:o + HID * HID] = rng.normal(0, np.sqrt(2.0 / HID), HID * HID); o += HID * HID
    p[o:o + HID] = 0.01; o += HID
    p[o:o + HID] = rng.normal(0, np.sqrt(1.0 / HID), HID); o += HID
    return p


def rank_table(ctx, params, width, height, colours):
    """uint16 [P, T] visit ranks (T = 4 x pieces of the largest kind): at each
    strategy index the transitions of the cell's kind by decreasing score
    (ties by transition), then the unused entries.  The histogram input at
    index k is the expected one (expected_histogram)."""
    colours = np.asarray(colours).reshape(-1, 4)
    P = width * height
    _, by_kind = piece_kinds(colours)
    T = 4 * max(len(b) for b in by_kind)
    order_xy = emp.order(width, height)
    feats, spans = [], []
    for k, (x, y) in enumerate(order_xy):
        K = cell_kind(x, y, width, height)
        cands = [(int(p), o) for p in by_kind[K] for o in range(4)]
        spans.append((K, len(cands)))
        feats.append(features(colours, P, k, cands, expected_histogram(colours, width, height, k)))
    scores = ctx.mlp_score(params, np.concatenate(feats))
    ranks = np.zeros((P, T), dtype=np.uint16)
    at = 0
    for k, (K, n) in enumerate(spans):
        s = scores[at:at + n]
        at += n
        order = np.lexsort((np.arange(n), -s))      # by score, then transition
        ranks[k, order] = np.arange(n, dtype=np.uint16)
        ranks[k, n:] = np.arange(n, T, dtype=np.uint16)
    return ranks


def relabel_by(colours, priority):
    """The puzzle with its pieces renumbered by decreasing priority (ties by
    id), except that the first corner piece stays the first corner: index 0
    only ever takes its first candidate (include/emp_strategy.h:389), so that
    piece fixes the tree searched, and keeping it leaves the tree -- node
    count and solutions up to identical pieces -- unchanged; only the visit
    order changes.  Returns (colours in the new numbering, old id of each new
    id)."""
    colours = np.asarray(colours).reshape(-1, 4)
    kind, by_kind = piece_kinds(colours)
    prio = np.asarray(priority, dtype=np.float64).copy()
    prio[by_kind[2][0]] = np.inf
    perm = np.lexsort((np.arange(len(prio)), -prio))
    return colours[perm].copy(), perm


def shuffled(colours, seed):
    """A random renumbering that keeps the tree (relabel_by)."""
    colours = np.asarray(colours).reshape(-1, 4)
    return relabel_by(colours, np.random.default_rng(seed).random(len(colours)))


def relabelled(ctx, params, width, height, colours):
    """The puzzle with its pieces renumbered by decreasing score, so that the
    plain order (kind index = id order within a kind) visits the better
    scored pieces first at full speed.  A piece's score is its best
    orientation's score averaged over the strategy indices of its kind; the
    first corner piece keeps its place (relabel_by).  Returns (colours in the
    new numbering, old id of each new id)."""
    colours = np.asarray(colours).reshape(-1, 4)
    P = width * height
    _, by_kind = piece_kinds(colours)
    order_xy = emp.order(width, height)
    prio = np.zeros(P, dtype=np.float64)
    for K in range(3):
        ks = [k for k, (x, y) in enumerate(order_xy) if cell_kind(x, y, width, height) == K]
        if not ks or len(by_kind[K]) == 0:
            continue
        cands = [(int(p), o) for p in by_kind[K] for o in range(4)]
        feats = np.concatenate([features(colours, P, k, cands, expected_histogram(colours, width, height, k))
                                for k in ks])
        sc = ctx.mlp_score(params, feats).reshape(len(ks), len(by_kind[K]), 4)
        for i, p in enumerate(by_kind[K]):
            kept = list(orientations_kept(colours, p))
            prio[p] = sc[:, i, kept].max(axis=1).mean()
    return relabel_by(colours, prio)


def random_rank_table(width, height, colours, seed):
    """A random visit order per strategy index (for exactness checks)."""
    colours = np.asarray(colours).reshape(-1, 4)
    _, by_kind = piece_kinds(colours)
    T = 4 * max(len(b) for b in by_kind)
    rng = np.random.default_rng(seed)
    ranks = np.zeros((width * height, T), dtype=np.uint16)
    for k, (x, y) in enumerate(emp.order(width, height)):
        n = 4 * len(by_kind[cell_kind(x, y, width, height)])
        ranks[k, :n] = rng.permutation(n)
        ranks[k, n:] = np.arange(n, T)
    return ranks


def experiment(ctx, situations_per_seed=64, probe_nodes=20000, epochs=10, batch=256, lr=1e-3, time_ms=2000.0,
               train_seeds=None, heldout_seeds=None, log=None):
    """Train the scorer on depth-probe traces of the training seeds and
    compare the max depth reached on the held-out seeds, for the same device
    time, by the plain order, the learned order and a random order."""
    import time
    train_seeds = TRAIN_SEEDS if train_seeds is None else train_seeds
    heldout_seeds = HELDOUT_SEEDS if heldout_seeds is None else heldout_seeds
    say = log or (lambda m: None)
    t0 = time.perf_counter()
    Xs, ys = [], []
    for seed in train_seeds:
        w, h, c = puzzle(seed, **C3_SHAPE)
        X, y, _ = traces(ctx, w, h, c, situations_per_seed, seed, probe_nodes, depths=(4, 180))
        Xs.append(X)
        ys.append(y)
    X, y = np.concatenate(Xs), np.concatenate(ys)
    trace_s = time.perf_counter() - t0
    say("traces: %d rows from %d seeds in %.1f s" % (len(y), len(train_seeds), trace_s))
    t0 = time.perf_counter()
    params, losses = ctx.mlp_train(X, y, init_params(1), epochs=epochs, batch=batch, lr=lr, seed=1)
    train_s = time.perf_counter() - t0
    say("training: %d epochs in %.1f s, loss %.4f -> %.4f" % (epochs, train_s, losses[0], losses[-1]))
    # held-out trace error against predicting the training mean
    hx, hy = [], []
    for seed in heldout_seeds[:4]:
        w, h, c = puzzle(seed, **C3_SHAPE)
        Xh, yh, _ = traces(ctx, w, h, c, 16, seed + 7, probe_nodes, depths=(4, 180))
        hx.append(Xh)
        hy.append(yh)
    Xh, yh = np.concatenate(hx), np.concatenate(hy)
    pred = ctx.mlp_score(params, Xh)
    heldout_mse = float(np.mean((pred - yh) ** 2))
    mean_mse = float(np.mean((y.mean() - yh) ** 2))
    rows = []
    for seed in heldout_seeds:
        w, h, c = puzzle(seed, **C3_SHAPE)
        t0 = time.perf_counter()
        ranks = rank_table(ctx, params, w, h, c)
        rank_s = time.perf_counter() - t0
        out = {"seed": seed, "rank_table_s": rank_s}
        relab, _ = relabelled(ctx, params, w, h, c)
        shuf, _ = shuffled(c, seed)
        for name, cc, rk in (("plain", c, None), ("learned", c, [ranks]), ("random", c, [random_rank_table(w, h, c, seed)]),
                             ("learned_relabel", relab, None), ("random_relabel", shuf, None)):
            (r,), st = ctx.enumerate([(w, h, cc)], max_solutions=0, ranks=rk, time_limit_ms=time_ms)
            out[name] = {"max_depth": r["max_depth"], "nodes": r["n_nodes"], "kernel_ms": st["dfs_kernel_ms"]}
        rows.append(out)
        say("seed %d: plain %d, learned %d, random %d, learned_relabel %d, random_relabel %d" % (
            seed, out["plain"]["max_depth"], out["learned"]["max_depth"], out["random"]["max_depth"],
            out["learned_relabel"]["max_depth"], out["random_relabel"]["max_depth"]))

    def summary(name):
        d = np.array([r[name]["max_depth"] for r in rows], dtype=float)
        n = np.array([r[name]["nodes"] for r in rows], dtype=float)
        ms = np.array([r[name]["kernel_ms"] for r in rows], dtype=float)
        return {"mean_max_depth": float(d.mean()), "max_max_depth": int(d.max()),
                "nodes_per_s": float(n.sum() / (ms.sum() * 1e-3))}

    return {"trace_rows": int(len(y)), "trace_s": trace_s, "train_s": train_s, "loss_first": float(losses[0]),
            "loss_last": float(losses[-1]), "heldout_mse": heldout_mse, "heldout_mse_of_mean": mean_mse,
            "time_ms_per_run": time_ms, "plain": summary("plain"), "learned": summary("learned"),
            "random": summary("random"), "learned_relabel": summary("learned_relabel"),
            "random_relabel": summary("random_relabel"),
            "relabel_deeper": int(sum(r["learned_relabel"]["max_depth"] > r["plain"]["max_depth"] for r in rows)),
            "relabel_shallower": int(sum(r["learned_relabel"]["max_depth"] < r["plain"]["max_depth"] for r in rows)),
            "learned_deeper": int(sum(r["learned"]["max_depth"] > r["plain"]["max_depth"] for r in rows)),
            "learned_shallower": int(sum(r["learned"]["max_depth"] < r["plain"]["max_depth"] for r in rows)),
            "per_seed": rows, "params": params}


def situation_rows(ctx, width, height, colours, depth, capacity):
    """Every situation of the enumeration with `depth` pieces placed (in the
    DFS's order), as packed moves [n, depth]; None when there are more than
    `capacity`."""
    try:
        boards, n, _, _ = ctx.expand(width, height, colours, depth, capacity=capacity)
    except emp.EmpError:
        return None   # a level did not fit the device queues
    if n > capacity:
        return None
    xy = emp.order(width, height)
    cells = np.array([y * width + x for x, y in xy[:depth]], dtype=np.uint32)
    b = boards[:, cells].astype(np.uint32)
    return cells[None, :] | ((b >> 2) << 8) | ((b & 3) << 16)


def row_features(colours, width, height, rows):
    """Features of each row's last placement in its situation (vectorised
    form of `features`)."""
    colours = np.asarray(colours).reshape(-1, 4).astype(np.int64)
    P = width * height
    n, d = rows.shape
    pieces = ((rows >> 8) & 0xFF).astype(np.int64)
    orient = ((rows >> 16) & 3).astype(np.int64)
    onehot = np.zeros((P, N_COLOURS), dtype=np.float32)
    for p in range(P):
        for col in colours[p]:
            if 1 <= col <= N_COLOURS:
                onehot[p, col - 1] += 1
    remaining = onehot.sum(0)[None, :] - onehot[pieces[:, :-1]].sum(1)   # before the last placement
    X = np.zeros((n, N_FEATURES), dtype=np.float32)
    X[:, 4 * N_COLOURS:5 * N_COLOURS] = remaining / (4.0 * (P - (d - 1)))
    X[:, 5 * N_COLOURS:] = position_embedding(d - 1, P)[None, :]
    p, o = pieces[:, -1], orient[:, -1]
    for s in range(4):
        col = colours[p, (s + o) & 3]
        ok = (col >= 1) & (col <= N_COLOURS)
        X[np.flatnonzero(ok), s * N_COLOURS + col[ok] - 1] = 1.0
    return X


def item_order_experiment(ctx, params, seeds, depth=6, capacity=1 << 18, time_ms=2000.0, log=None):
    """Max depth reached in `time_ms` of device time when the search starts
    from the situations of one depth taken in DFS order (plain), by
    decreasing score of their last placement (learned), or shuffled
    (random)."""
    say = log or (lambda m: None)
    rows_out = []
    for seed in seeds:
        w, h, c = puzzle(seed, **C3_SHAPE)
        d = depth
        rows = situation_rows(ctx, w, h, c, d, capacity)
        while rows is None and d > 2:
            d -= 1
            rows = situation_rows(ctx, w, h, c, d, capacity)
        scores = ctx.mlp_score(params, row_features(c, w, h, rows))
        orders = {"plain": np.arange(len(rows)), "learned": np.lexsort((np.arange(len(rows)), -scores)),
                  "random": np.random.default_rng(seed).permutation(len(rows))}
        out = {"seed": seed, "depth": d, "situations": int(len(rows))}
        for name, perm in orders.items():
            res, st = ctx.probe(w, h, c, rows[perm], node_limit=0, time_limit_ms=time_ms)
            out[name] = {"max_depth": max(r["max_depth"] for r in res), "nodes": int(sum(r["n_nodes"] for r in res)),
                         "kernel_ms": st["dfs_kernel_ms"]}
        rows_out.append(out)
        say("seed %d (%d situations at depth %d): plain %d, learned %d, random %d" % (
            seed, len(rows), d, out["plain"]["max_depth"], out["learned"]["max_depth"], out["random"]["max_depth"]))
    summ = {}
    for name in ("plain", "learned", "random"):
        dd = np.array([r[name]["max_depth"] for r in rows_out], dtype=float)
        summ[name] = {"mean_max_depth": float(dd.mean()), "max_max_depth": int(dd.max())}
    summ["learned_deeper"] = int(sum(r["learned"]["max_depth"] > r["plain"]["max_depth"] for r in rows_out))
    summ["learned_shallower"] = int(sum(r["learned"]["max_depth"] < r["plain"]["max_depth"] for r in rows_out))
    return summ, rows_out
