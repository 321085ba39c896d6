"""Validity of a (partial) board, checked from the puzzle alone.

Independent of the oracle and of the library: it restates only the edge
rules of the reference's model -- an oriented piece (id, o) shows colour
colours[id - 1][(s + o) % 4] on side s (include/emp_piece.h:173-178),
colour 0 is the frame and appears exactly on the sides that face it, and two
placed neighbours show the same colour on their shared edge.

A board is row-major, one uint16 per cell: ((id - 1) << 2) | o, 0xFFFF empty.
"""
import numpy as np

EMPTY = 0xFFFF
# neighbour offset (dx, dy) of side s = N, E, S, W
_SIDES = ((0, -1), (1, 0), (0, 1), (-1, 0))


def board_problems(w, h, colours, board, prefix_moves=(), depth=None, solved=None):
    """List of rule violations of `board` (empty when valid).

    prefix_moves: packed moves (pos | piece << 8 | o << 16) that must be on
    the board as given (the first min(L, pieces placed) of them: a search
    whose prefix dies keeps only the moves it could play).  depth: the
    number of placed pieces it must hold.  solved: True requires a full
    board, False a non-full one."""
    colours = np.asarray(colours)
    board = np.asarray(board, dtype=np.int64).reshape(-1)
    P = w * h
    out = []
    if board.shape[0] != P:
        return ["board has %d cells, expected %d" % (board.shape[0], P)]
    placed = board != EMPTY
    ids = board[placed] >> 2
    if (ids >= P).any():
        out.append("piece index out of range")
        return out
    uniq, cnt = np.unique(ids, return_counts=True)
    for q in uniq[cnt > 1]:
        out.append("piece %d placed %d times" % (q + 1, int(cnt[uniq == q][0])))

    def shown(cell, side):
        v = int(board[cell])
        return int(colours[v >> 2][(side + (v & 3)) % 4])

    for cell in range(P):
        if not placed[cell]:
            continue
        x, y = cell % w, cell // w
        for s, (dx, dy) in enumerate(_SIDES):
            nx, ny = x + dx, y + dy
            c = shown(cell, s)
            if not (0 <= nx < w and 0 <= ny < h):
                if c != 0:
                    out.append("cell (%d,%d) shows colour %d towards the frame on side %d" % (x, y, c, s))
                continue
            if c == 0:
                out.append("cell (%d,%d) shows the frame colour inside the board on side %d" % (x, y, s))
                continue
            ncell = ny * w + nx
            if placed[ncell] and s < 2:   # each shared edge once (N/E of this cell)
                cn = shown(ncell, (s + 2) % 4)
                if cn != c:
                    out.append("edge between (%d,%d) and (%d,%d): colours %d and %d" % (x, y, nx, ny, c, cn))
    # The deepest situation starts with the prefix moves in order; a prefix
    # whose last moves could not be played leaves only its first ones.
    n_placed = int(placed.sum())
    for m in list(prefix_moves)[:n_placed]:
        m = int(m)
        pos, piece, o = m & 0xFF, (m >> 8) & 0xFF, (m >> 16) & 3
        if pos >= P or int(board[pos]) != (piece << 2 | o):
            out.append("prefix move %d (piece %d, orientation %d) not on the board" % (pos, piece + 1, o))
    n = n_placed
    if depth is not None and n != depth:
        out.append("board holds %d pieces, depth is %d" % (n, depth))
    if solved is True and n != P:
        out.append("solved but %d of %d cells placed" % (n, P))
    if solved is False and n == P:
        out.append("full board reported as not solved")
    return out


def assert_valid(w, h, colours, board, **kw):
    problems = board_problems(w, h, colours, board, **kw)
    assert not problems, "; ".join(problems[:8])


def invalid_boards(w, h, colours, boards, prefixes=None, depths=None, solved=None):
    """Vectorised form of board_problems over boards [n, P]: the indices of
    the boards that break a rule.  prefixes [n, L] packed moves, depths [n],
    solved [n] bools, each optional."""
    colours = np.asarray(colours, dtype=np.int64)
    b = np.asarray(boards, dtype=np.int64).reshape(-1, w * h)
    n, P = b.shape
    placed = b != EMPTY
    bad = np.zeros(n, dtype=bool)
    piece = np.where(placed, b >> 2, 0)
    bad |= (piece >= P).any(axis=1)
    piece = np.minimum(piece, P - 1)
    o = b & 3
    # shown[n, P, s] = colour on side s
    side = np.arange(4)[None, None, :]
    shown = colours[piece[:, :, None], (side + o[:, :, None]) % 4]
    xs, ys = np.arange(P) % w, np.arange(P) // w
    for s, (dx, dy) in enumerate(_SIDES):
        nx, ny = xs + dx, ys + dy
        out_edge = (nx < 0) | (nx >= w) | (ny < 0) | (ny >= h)
        c = shown[:, :, s]
        # frame sides show 0, inner sides never do
        bad |= (placed & out_edge[None, :] & (c != 0)).any(axis=1)
        bad |= (placed & ~out_edge[None, :] & (c == 0)).any(axis=1)
        inside = np.nonzero(~out_edge)[0]
        ncell = ny[inside] * w + nx[inside]
        both = placed[:, inside] & placed[:, ncell]
        cn = shown[:, ncell, (s + 2) % 4]
        bad |= (both & (c[:, inside] != cn)).any(axis=1)
    # each piece at most once
    counts = np.zeros((n, P), dtype=np.int64)
    rows = np.repeat(np.arange(n), P)
    np.add.at(counts, (rows[placed.reshape(-1)], piece.reshape(-1)[placed.reshape(-1)]), 1)
    bad |= (counts > 1).any(axis=1)
    npl = placed.sum(axis=1)
    if prefixes is not None:
        pre = np.asarray(prefixes, dtype=np.int64).reshape(n, -1)
        pos = pre & 0xFF
        want = ((pre >> 8) & 0xFF) << 2 | ((pre >> 16) & 3)
        bad |= (pos >= P).any(axis=1)
        got = np.take_along_axis(b, np.minimum(pos, P - 1), axis=1)
        kept = np.arange(pre.shape[1])[None, :] < np.minimum(placed.sum(axis=1), pre.shape[1])[:, None]
        bad |= (kept & (got != want)).any(axis=1)
    if depths is not None:
        bad |= npl != np.asarray(depths).reshape(n)
    if solved is not None:
        sv = np.asarray(solved, dtype=bool).reshape(n)
        bad |= sv != (npl == P)
    return np.nonzero(bad)[0]
