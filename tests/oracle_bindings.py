"""ctypes bindings of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: the oracle is the checker, never the product.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

OK, BAD_ARGUMENT, BAD_PUZZLE, LIMIT_REACHED = 0, 1, 2, 3
SPIRAL, BASIC = 0, 1

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _lib = ctypes.CDLL(LIB_PATH)
        u32, u64, i32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
        p = ctypes.c_void_p
        _lib.oracle_enumerate.argtypes = [u32, u32, p, i32, u64, u64, p, p, p, p, p, p, p, u64]
        _lib.oracle_enumerate.restype = i32
        _lib.oracle_results_bin.argtypes = [u32, u32, p, i32, p, u64, p]
        _lib.oracle_results_bin.restype = i32
        _lib.oracle_order.argtypes = [u32, u32, i32, p]
        _lib.oracle_order.restype = i32
        _lib.oracle_enumerate_prefixes.argtypes = [u32, u32, p, i32, u32, p, u64, p, p, p]
        _lib.oracle_enumerate_prefixes.restype = i32
        _lib.oracle_enumerate_subtree.argtypes = [u32, u32, p, i32, u32, p, u64, p, p, p]
        _lib.oracle_enumerate_subtree.restype = i32
        _lib.oracle_enumerate_resume.argtypes = [u32, u32, p, i32, p, u64, p, p, p, p]
        _lib.oracle_enumerate_resume.restype = i32
        _lib.oracle_subtree_depth.argtypes = [u32, u32, p, i32, u32, p, p, p]
        _lib.oracle_subtree_depth.restype = i32
        _lib.oracle_backtracker.argtypes = [u32, u32, p, i32, p, p, p, u64, p]
        _lib.oracle_backtracker.restype = i32
        _lib.oracle_glutton_masks.argtypes = [u32, u32, p, u32, p, p, p]
        _lib.oracle_glutton_masks.restype = i32
        _lib.oracle_glutton_max.argtypes = [u32, u32, p, u32, p, u64, p, p, p, p]
        _lib.oracle_glutton_max.restype = i32
        _lib.oracle_situation_profile.argtypes = [u32, u32, p, p, p, p, p, p, p]
        _lib.oracle_situation_profile.restype = i32
        _lib.oracle_profile.argtypes = [u32, u32, p, u64, p, p, u32, p]
        _lib.oracle_profile.restype = i32
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def parse_puzzle(path):
    """Reference text format: 'Width,Height' / 'W,H' / 'I,N,E,S,O' / pieces."""
    lines = [l.strip() for l in open(path) if l.strip()]
    w, h = (int(v) for v in lines[1].split(","))
    rows = [[int(v) for v in l.split(",")] for l in lines[3:]]
    rows.sort(key=lambda r: r[0])
    colours = np.array([r[1:5] for r in rows], dtype=np.uint32)
    assert colours.shape == (w * h, 4)
    return w, h, colours


def enumerate_(w, h, colours, order=SPIRAL, node_limit=0, max_solutions=1 << 16, trace=False):
    colours = np.ascontiguousarray(colours, dtype=np.uint32)
    P = w * h
    boards = np.zeros((max_solutions, P), dtype=np.uint16)
    snodes = np.zeros(max_solutions, dtype=np.uint64)
    nsol = np.zeros(1, dtype=np.uint64)
    nnodes = np.zeros(1, dtype=np.uint64)
    dn = np.zeros(P, dtype=np.uint64)
    dd = np.zeros(P, dtype=np.uint64)
    cap = 1 << 20 if trace else 0
    buf = ctypes.create_string_buffer(cap) if trace else None
    st = lib().oracle_enumerate(w, h, _ptr(colours), order, node_limit, max_solutions,
                                _ptr(boards), _ptr(snodes), _ptr(nsol), _ptr(nnodes),
                                _ptr(dn), _ptr(dd), buf, cap)
    if st not in (OK, LIMIT_REACHED):
        raise ValueError("oracle_enumerate status %d" % st)
    n = int(nsol[0])
    kept = min(n, max_solutions)
    return {
        "status": st, "n_solutions": n, "n_nodes": int(nnodes[0]),
        "boards": boards[:kept].copy(), "solution_nodes": snodes[:kept].copy(),
        "depth_nodes": dn, "depth_dead": dd,
        "trace": buf.value.decode() if trace else None,
    }


def prefixes(w, h, colours, depth, order=SPIRAL, capacity=1 << 22):
    """Placements at index depth-1 of the DFS; returns (prefixes [n, depth],
    nodes above the cut, solutions above the cut)."""
    colours = np.ascontiguousarray(colours, dtype=np.uint32)
    out = np.zeros((capacity, depth), dtype=np.uint16)
    n, nodes, sols = (np.zeros(1, np.uint64) for _ in range(3))
    st = lib().oracle_enumerate_prefixes(w, h, _ptr(colours), order, depth, _ptr(out), capacity,
                                         _ptr(n), _ptr(nodes), _ptr(sols))
    if st != OK:
        raise ValueError("oracle_enumerate_prefixes status %d" % st)
    return out[:int(n[0])].copy(), int(nodes[0]), int(sols[0])


def subtree(w, h, colours, prefix, order=SPIRAL, max_solutions=1 << 12):
    colours = np.ascontiguousarray(colours, dtype=np.uint32)
    prefix = np.ascontiguousarray(prefix, dtype=np.uint16)
    boards = np.zeros((max_solutions, w * h), dtype=np.uint16)
    sols, nodes = np.zeros(1, np.uint64), np.zeros(1, np.uint64)
    st = lib().oracle_enumerate_subtree(w, h, _ptr(colours), order, prefix.size, _ptr(prefix),
                                        max_solutions, _ptr(boards), _ptr(sols), _ptr(nodes))
    if st != OK:
        raise ValueError("oracle_enumerate_subtree status %d" % st)
    n = int(sols[0])
    return {"n_nodes": int(nodes[0]), "n_solutions": n, "boards": boards[:min(n, max_solutions)].copy()}


def resume(w, h, colours, situation, order=SPIRAL, max_solutions=1 << 12):
    """The reference's search resumed from a situation (row-major board)."""
    colours = np.ascontiguousarray(colours, dtype=np.uint32)
    sit = np.ascontiguousarray(situation, dtype=np.uint16)
    boards = np.zeros((max_solutions, w * h), dtype=np.uint16)
    sol_nodes = np.zeros(max_solutions, dtype=np.uint64)
    ns, nn = np.zeros(1, np.uint64), np.zeros(1, np.uint64)
    st = lib().oracle_enumerate_resume(w, h, _ptr(colours), order, _ptr(sit), max_solutions, _ptr(boards),
                                       _ptr(sol_nodes), _ptr(ns), _ptr(nn))
    if st != OK:
        raise ValueError("oracle_enumerate_resume status %d" % st)
    k = min(int(ns[0]), max_solutions)
    return {"n_nodes": int(nn[0]), "n_solutions": int(ns[0]), "boards": boards[:k].copy(),
            "solution_nodes": sol_nodes[:k].copy()}


def subtree_depth(w, h, colours, prefix, order=SPIRAL):
    """(placements below the prefix, most pieces placed in one situation)."""
    colours = np.ascontiguousarray(colours, dtype=np.uint32)
    prefix = np.ascontiguousarray(prefix, dtype=np.uint16)
    nodes, depth = np.zeros(1, np.uint64), np.zeros(1, np.uint32)
    st = lib().oracle_subtree_depth(w, h, _ptr(colours), order, prefix.size, _ptr(prefix), _ptr(nodes), _ptr(depth))
    if st != OK:
        raise ValueError("oracle_subtree_depth status %d" % st)
    return int(nodes[0]), int(depth[0])


def results_bin(w, h, colours, order=SPIRAL):
    colours = np.ascontiguousarray(colours, dtype=np.uint32)
    size = np.zeros(1, dtype=np.uint64)
    st = lib().oracle_results_bin(w, h, _ptr(colours), order, None, 0, _ptr(size))
    if st != OK:
        raise ValueError("oracle_results_bin status %d" % st)
    out = np.zeros(int(size[0]), dtype=np.uint8)
    lib().oracle_results_bin(w, h, _ptr(colours), order, _ptr(out), out.size, _ptr(size))
    return out.tobytes()


def order(w, h, kind=SPIRAL):
    xy = np.zeros(2 * w * h, dtype=np.uint32)
    lib().oracle_order(w, h, kind, _ptr(xy))
    return [(int(xy[2 * i]), int(xy[2 * i + 1])) for i in range(w * h)]


def glutton_masks(w, h, colours, moves=()):
    """Position masks [P, nw] u64 after the (pos, piece, o) moves and the
    status (0 valid, 1 a position emptied, 2 unplayable move)."""
    colours = np.ascontiguousarray(colours, dtype=np.uint32)
    mv = np.ascontiguousarray(np.array(list(moves), dtype=np.uint32).reshape(-1, 3))
    nw = (4 * w * h + 63) // 64
    masks = np.zeros((w * h, nw), dtype=np.uint64)
    status = np.zeros(1, dtype=np.uint32)
    st = lib().oracle_glutton_masks(w, h, _ptr(colours), mv.shape[0], _ptr(mv), _ptr(masks), _ptr(status))
    if st != OK:
        raise ValueError("oracle_glutton_masks status %d" % st)
    return masks, int(status[0])


def glutton_max(w, h, colours, prefix=(), push_limit=0):
    colours = np.ascontiguousarray(colours, dtype=np.uint32)
    moves = np.array(list(prefix), dtype=np.uint32).reshape(-1, 3) if len(prefix) else np.zeros((0, 3), np.uint32)
    moves = np.ascontiguousarray(moves)
    board = np.zeros(w * h, dtype=np.uint16)
    md = np.zeros(1, dtype=np.uint32)
    pushes = np.zeros(1, dtype=np.uint64)
    solved = np.zeros(1, dtype=np.uint32)
    st = lib().oracle_glutton_max(w, h, _ptr(colours), moves.shape[0], _ptr(moves), push_limit,
                                  _ptr(board), _ptr(md), _ptr(pushes), _ptr(solved))
    if st not in (OK, LIMIT_REACHED):
        raise ValueError("oracle_glutton_max status %d" % st)
    return {"status": st, "board": board, "max_depth": int(md[0]), "pushes": int(pushes[0]),
            "solved": bool(solved[0])}


def backtracker(w, h, colours, order=BASIC, trace_capacity=1 << 20):
    colours = np.ascontiguousarray(colours, dtype=np.uint32)
    board = np.zeros(w * h, dtype=np.uint16)
    found = np.zeros(1, np.uint32)
    trace = np.zeros((trace_capacity, 3), np.uint32)
    n = np.zeros(1, np.uint64)
    st = lib().oracle_backtracker(w, h, _ptr(colours), order, _ptr(board), _ptr(found), _ptr(trace),
                                  trace_capacity, _ptr(n))
    if st != OK:
        raise ValueError("oracle_backtracker status %d" % st)
    return {"found": bool(found[0]), "board": board, "trace": trace[:min(int(n[0]), trace_capacity)].copy(),
            "trace_length": int(n[0])}


def situation_string(board, P):
    w = len(str(P)) + 1
    out = []
    for v in board:
        v = int(v)
        if v == 0xFFFF:
            out.append("-" * w)
        else:
            out.append(str((v >> 2) + 1).rjust(w - 1) + "NESW"[v & 3])
    return "".join(out)


def situation_profile(w, h, colours, board):
    """Capability profiles of a situation (row-major board, 0xFFFF empty)."""
    colours = np.ascontiguousarray(colours, dtype=np.uint32)
    board = np.ascontiguousarray(board, dtype=np.uint16)
    P = w * h
    level = np.zeros(1, np.uint32)
    glutton = np.zeros((10, P + 1), np.uint32)
    surface = np.zeros((P + 1, 2 * P), np.uint32)
    row_index = np.zeros(P + 1, np.uint32)
    nvar = np.zeros(1, np.uint32)
    st = lib().oracle_situation_profile(w, h, _ptr(colours), _ptr(board), _ptr(level), _ptr(glutton), _ptr(surface),
                                        _ptr(row_index), _ptr(nvar))
    if st != OK:
        raise ValueError("oracle_situation_profile status %d" % st)
    L = int(level[0])
    return {"level": L, "glutton": glutton, "surface": surface[:L + 1].copy(), "row_index": row_index[:L + 1].copy(),
            "n_variables": int(nvar[0])}


PROFILE_STATS = ("low_min", "low_max", "high_min", "high_max", "total_min", "total_max")


def profile(w, h, colours, max_situations=1 << 22, capacity=4096):
    """The profile feature: per-level extrema [6 stats][3 series][P + 1]
    (PROFILE_STATS; series all / leading to a solution / others) and the
    solution set (boards, ordered)."""
    colours = np.ascontiguousarray(colours, dtype=np.uint32)
    P = w * h
    stats = np.zeros((6, 3, P + 1), np.uint32)
    sols = np.zeros((capacity, P), np.uint16)
    n = np.zeros(1, np.uint32)
    st = lib().oracle_profile(w, h, _ptr(colours), max_situations, _ptr(stats), _ptr(sols), capacity, _ptr(n))
    if st != OK:
        raise ValueError("oracle_profile status %d" % st)
    return {"stats": stats, "solutions": sols[:min(int(n[0]), capacity)].copy(), "n_solutions": int(n[0])}
