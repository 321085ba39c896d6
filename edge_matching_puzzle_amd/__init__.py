"""edge_matching_puzzle_amd: MI355X (gfx950) search path of edge_matching_puzzle.

Python binding of libemp_hip.so, the C ABI declared in include/emp_hip.h.
The library is the product: every search runs in its HIP kernels.  There is
no CPU fallback -- loading fails loudly when the library is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libemp_hip.so")

OK, BAD_ARGUMENT, BAD_PUZZLE, DEVICE_ERROR, OUT_OF_MEMORY, UNSUPPORTED = range(6)
ORDER_SPIRAL, ORDER_BASIC = 0, 1
_ORDERS = {"spiral": ORDER_SPIRAL, "basic": ORDER_BASIC}


class EmpError(RuntimeError):
    """Error status returned by libemp_hip.so."""

    def __init__(self, status, message):
        super().__init__("emp_hip error %d: %s" % (status, message))
        self.status = status


class Puzzle(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("colours", ctypes.c_void_p)]


# numpy records laid out as emp_hip_puzzle and emp_hip_enum_result, for batches
_PUZZLE_RECORD = np.dtype([("width", "<u4"), ("height", "<u4"), ("colours", "<u8")])
assert _PUZZLE_RECORD.itemsize == ctypes.sizeof(Puzzle)


class EnumOptions(ctypes.Structure):
    _fields_ = [("order", ctypes.c_uint32), ("waves_per_cu", ctypes.c_uint32),
                ("waves_per_block", ctypes.c_uint32), ("steps_per_round", ctypes.c_uint32),
                ("step_budget", ctypes.c_uint64), ("target_items", ctypes.c_uint64),
                ("node_limit", ctypes.c_uint64), ("max_solutions", ctypes.c_uint64),
                ("use_graph", ctypes.c_uint32), ("queue_items", ctypes.c_uint32),
                ("time_limit_ms", ctypes.c_double), ("count_work", ctypes.c_uint32),
                ("schedule", ctypes.c_uint32)]


class EnumResult(ctypes.Structure):
    _fields_ = [("n_nodes", ctypes.c_uint64), ("n_solutions", ctypes.c_uint64),
                ("n_stored", ctypes.c_uint64), ("complete", ctypes.c_uint32),
                ("status", ctypes.c_uint32), ("solve_ms", ctypes.c_double),
                ("max_depth", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


_ENUM_RESULT_RECORD = np.dtype([(f, {ctypes.c_uint64: "<u8", ctypes.c_uint32: "<u4", ctypes.c_double: "<f8"}[t])
                                for f, t in EnumResult._fields_])
assert _ENUM_RESULT_RECORD.itemsize == ctypes.sizeof(EnumResult)


class EnumStats(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_double), ("dfs_kernel_ms", ctypes.c_double),
                ("expand_kernel_ms", ctypes.c_double), ("dfs_launches", ctypes.c_uint64),
                ("expand_launches", ctypes.c_uint64), ("items_processed", ctypes.c_uint64),
                ("splits", ctypes.c_uint64), ("max_words", ctypes.c_uint32),
                ("lds_tables", ctypes.c_uint32), ("blocks", ctypes.c_uint32),
                ("threads", ctypes.c_uint32), ("candidate_sets", ctypes.c_uint64),
                ("group_ms", ctypes.c_double), ("work_bytes", ctypes.c_uint64),
                ("work_ops", ctypes.c_uint64), ("tickets", ctypes.c_uint64)]


class GluttonOptions(ctypes.Structure):
    _fields_ = [("push_budget", ctypes.c_uint64), ("waves_per_cu", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class GluttonResult(ctypes.Structure):
    _fields_ = [("max_depth", ctypes.c_uint32), ("status", ctypes.c_uint32), ("pushes", ctypes.c_uint64)]


class GluttonStats(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_double), ("kernel_ms", ctypes.c_double),
                ("pushes", ctypes.c_uint64), ("evaluations", ctypes.c_uint64),
                ("waves", ctypes.c_uint32), ("lds_bytes", ctypes.c_uint32),
                ("score_evaluations", ctypes.c_uint64), ("rounds", ctypes.c_uint64),
                ("free_slot_sum", ctypes.c_uint64)]


GLUTTON_SOLVED, GLUTTON_EXHAUSTED, GLUTTON_BUDGET, GLUTTON_BAD_PREFIX = range(4)

NO_MOVE = 0xFFFFFFFF
# emp_hip_enum_options.schedule
SCHEDULES = {"auto": 0, "passes": 1, "persistent": 2}
WIDE_VALID, WIDE_EMPTIED, WIDE_BAD_MOVE = range(3)


class Item(ctypes.Structure):
    """emp_hip_item: a unit of remaining search (floor, level, last move)."""
    _fields_ = [("floor", ctypes.c_uint32), ("level", ctypes.c_uint32), ("last", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]

_lib = None

# Every symbol include/emp_hip.h declares, with (restype, argtypes).
_P = ctypes.c_void_p
_U32, _U64, _I32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
SIGNATURES = {
    "emp_hip_last_error": (ctypes.c_char_p, []),
    "emp_hip_version": (ctypes.c_char_p, []),
    "emp_hip_create": (_I32, [_I32, _P]),
    "emp_hip_destroy": (_I32, [_P]),
    "emp_hip_device_info": (_I32, [_P, _P, _P, _P]),
    "emp_hip_generate": (_I32, [_U64, _U32, _U32, _U32, _U32, _P, _P]),
    "emp_hip_validate_puzzle": (_I32, [_P]),
    "emp_hip_order": (_I32, [_U32, _U32, _U32, _P]),
    "emp_hip_situation_string": (_I32, [_U32, _U32, _P, ctypes.c_char_p, _U64]),
    "emp_hip_enum_default_options": (None, [_P]),
    "emp_hip_enumerate": (_I32, [_P, _P, _U32, _P, _P, _P, _P]),
    "emp_hip_enumerate_noalloc": (_I32, [_P, _P, _U32, _P, _P, _P, _P]),
    "emp_hip_enumerate_items": (_I32, [_P, _P, _P, _P, _P, _U32, _P, _P, _P, _P, _P, _U64, _P, _P]),
    "emp_hip_resume_item": (_I32, [_P, _U32, _P, _P, _P, _P]),
    "emp_hip_parse_situation": (_I32, [_U32, _U32, ctypes.c_char_p, _P]),
    "emp_hip_results_bin_write": (_I32, [_U32, _U32, _U32, _U32, _P, _P, _U64, _U64, _P, _U64, _P]),
    "emp_hip_results_bin_read": (_I32, [_P, _U64, _P, _P, _P, _P, _P, _P, _U64, _P, _P, _P]),
    "emp_hip_backtrack_trace": (_I32, [_P, _P, _U32, _U64, _P, _P, _P, _U64, _P]),
    "emp_hip_expand": (_I32, [_P, _P, _U32, _U32, _U64, _P, _P, _P, _P]),
    "emp_hip_glutton_max": (_I32, [_P, _P, _P, _U32, _U32, _P, _P, _P, _P]),
    "emp_hip_glutton_stack_words": (_I32, [_U32, _P, _P]),
    "emp_hip_glutton_max_stacks": (_I32, [_P, _P, _P, _U32, _U32, _P, _P, _P, _P, _P, _P, _P, _P]),
    "emp_hip_glutton_stack_to_xml": (_I32, [_P, _P, _P, _P, _U64, _P]),
    "emp_hip_glutton_stack_from_xml": (_I32, [_P, _P, _U64, _P, _P]),
    "emp_hip_spiral_prefixes": (_I32, [_P, _U32, _U32, _U64, _P, _U32, _P, _P]),
    "emp_hip_backtrack_variables": (_I32, [_P, _U32, _P]),
    "emp_hip_glutton_wide": (_I32, [_P, _P, _P, _U32, _U32, _P, _P, _P, _U64, _P, _P, _P, _P, _P]),
    "emp_hip_situation_profile": (_I32, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "emp_hip_profile": (_I32, [_P, _P, _U64, _P, _P, _U64, _P, _P, _P, _P]),
    "emp_hip_enumerate_ranked": (_I32, [_P, _P, _U32, _P, _P, _P, _P, _P]),
    "emp_hip_probe": (_I32, [_P, _P, _P, _U32, _U32, _P, _P, _P]),
    "emp_hip_stream_copy": (_I32, [_P, _U64, _P]),
    "emp_hip_peaks": (_I32, [_P, _P, _P]),
    "emp_hip_mlp_train": (_I32, [_P, _P, _P, _U64, _P, _P, _P]),
    "emp_hip_mlp_score": (_I32, [_P, _P, _P, _U64, _P]),
    "emp_hip_mlp_gradients": (_I32, [_P, _P, _P, _P, _U64, _P, _P]),
}

MLP_INPUTS = 126
MLP_PARAMS = 99073


class MlpOptions(ctypes.Structure):
    _fields_ = [("epochs", ctypes.c_uint32), ("batch", ctypes.c_uint32), ("lr", ctypes.c_float),
                ("reserved", ctypes.c_uint32), ("seed", ctypes.c_uint64)]


def lib():
    """Loads libemp_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libemp_hip.so is not built; run `make -C edge_matching_puzzle_amd` "
                               "(or __graft_entry__.build())")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def _check(status):
    if status != OK:
        raise EmpError(status, lib().emp_hip_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def parse_puzzle(path):
    """Reads the reference's text format (include/emp_pieces_parser.h:53-104):
    'Width,Height', 'W,H', 'I,N,E,S,O', then 'id,n,e,s,w' per piece."""
    lines = [l.strip() for l in open(path) if l.strip()]
    w, h = (int(v) for v in lines[1].split(","))
    rows = sorted(([int(v) for v in l.split(",")] for l in lines[3:]), key=lambda r: r[0])
    if [r[0] for r in rows] != list(range(1, w * h + 1)):
        raise ValueError("%s: piece ids must be 1..%d" % (path, w * h))
    return w, h, np.array([r[1:5] for r in rows], dtype=np.uint32)


def generate(seed, width, height, border_colours, interior_colours):
    """Seeded synthetic puzzle; returns (colours [P,4] uint32, planted board)."""
    P = width * height
    colours = np.zeros((P, 4), dtype=np.uint32)
    planted = np.zeros(P, dtype=np.uint16)
    _check(lib().emp_hip_generate(seed, width, height, border_colours, interior_colours,
                                  _ptr(colours), _ptr(planted)))
    return colours, planted


def validate(width, height, colours):
    """Raises EmpError(BAD_PUZZLE) when the reference would refuse the puzzle."""
    colours = np.ascontiguousarray(colours, dtype=np.uint32)
    pz = Puzzle(width, height, colours.ctypes.data)
    _check(lib().emp_hip_validate_puzzle(ctypes.byref(pz)))


def kind_counts(colours):
    """Pieces of each kind (centre, border, corner: the number of frame
    sides, include/emp_piece.h:126-131)."""
    zeros = (np.asarray(colours) == 0).sum(axis=1)
    return [int((zeros == k).sum()) for k in range(3)]


def pack_move(pos, piece, orientation):
    return pos | (piece << 8) | (orientation << 16)


def spiral_prefixes(width, height, colours, length, count, seed, fixed=()):
    """Random distinct placements of the first `length` spiral positions,
    after the `fixed` packed moves; returns [n, len(fixed) + length] uint32."""
    colours = np.ascontiguousarray(colours, dtype=np.uint32)
    pz = Puzzle(width, height, colours.ctypes.data)
    fx = np.ascontiguousarray(np.array(list(fixed), dtype=np.uint32))
    out = np.zeros((count, len(fixed) + length), dtype=np.uint32)
    n = ctypes.c_uint32()
    _check(lib().emp_hip_spiral_prefixes(ctypes.byref(pz), length, count, seed,
                                         _ptr(fx) if len(fixed) else None, len(fixed), _ptr(out),
                                         ctypes.byref(n)))
    return out[:n.value].copy()


def parse_situation(width, height, text):
    """Row-major board (uint16, 0xFFFF empty) of a situation string
    (include/situation_string_formatter.h:84-134)."""
    board = np.zeros(width * height, dtype=np.uint16)
    _check(lib().emp_hip_parse_situation(width, height, text.encode(), _ptr(board)))
    return board


def backtrack_variables(width, height, colours, order="basic"):
    """Variables of the backtracker search (create_variables; the reference
    prints "Nb variables: <n>")."""
    colours = np.ascontiguousarray(colours, dtype=np.uint32)
    pz = Puzzle(width, height, colours.ctypes.data)
    n = ctypes.c_uint64()
    _check(lib().emp_hip_backtrack_variables(ctypes.byref(pz), _ORDERS[order], ctypes.byref(n)))
    return n.value


def glutton_stack_words(n_pieces):
    """(state words, mask words) of one saved glutton stack."""
    sw, mw = ctypes.c_uint32(), ctypes.c_uint64()
    _check(lib().emp_hip_glutton_stack_words(n_pieces, ctypes.byref(sw), ctypes.byref(mw)))
    return sw.value, mw.value


def glutton_stack_to_xml(width, height, colours, state, masks):
    """XML text of one saved glutton stack (the reference converter's
    element layout)."""
    colours = np.ascontiguousarray(colours, dtype=np.uint32)
    pz = Puzzle(width, height, colours.ctypes.data)
    state = np.ascontiguousarray(state, dtype=np.uint32)
    masks = np.ascontiguousarray(masks, dtype=np.uint64)
    size = ctypes.c_uint64()
    _check(lib().emp_hip_glutton_stack_to_xml(ctypes.byref(pz), _ptr(state), _ptr(masks), None, 0, ctypes.byref(size)))
    buf = ctypes.create_string_buffer(size.value)
    _check(lib().emp_hip_glutton_stack_to_xml(ctypes.byref(pz), _ptr(state), _ptr(masks), buf, size.value,
                                              ctypes.byref(size)))
    return buf.raw[:size.value].decode()


def glutton_stack_from_xml(width, height, colours, xml):
    """(state, masks) of an XML glutton stack."""
    colours = np.ascontiguousarray(colours, dtype=np.uint32)
    pz = Puzzle(width, height, colours.ctypes.data)
    sw, mw = glutton_stack_words(width * height)
    state = np.zeros(sw, dtype=np.uint32)
    masks = np.zeros(mw, dtype=np.uint64)
    data = xml.encode() if isinstance(xml, str) else bytes(xml)
    _check(lib().emp_hip_glutton_stack_from_xml(ctypes.byref(pz), data, len(data), _ptr(state), _ptr(masks)))
    return state, masks


def resume_item(width, height, colours, board, order="spiral"):
    """The work item the reference resumes from a situation
    (include/emp_strategy.h:418-467): (items [n, 3] uint32 floor/level/last,
    moves [n, P] uint32) with n = 0 (the whole enumeration) or 1."""
    colours = np.ascontiguousarray(colours, dtype=np.uint32)
    pz = Puzzle(width, height, colours.ctypes.data)
    b = np.ascontiguousarray(board, dtype=np.uint16)
    it = Item()
    moves = np.zeros(width * height, dtype=np.uint32)
    n = ctypes.c_uint32()
    _check(lib().emp_hip_resume_item(ctypes.byref(pz), _ORDERS[order], _ptr(b), ctypes.byref(it), _ptr(moves),
                                     ctypes.byref(n)))
    if n.value == 0:
        return np.zeros((0, 3), dtype=np.uint32), np.zeros((0, width * height), dtype=np.uint32)
    return np.array([[it.floor, it.level, it.last]], dtype=np.uint32), moves[None, :].copy()


def results_bin_write(width, height, boards, node_index, total_nodes, order="spiral", solution_dump=True):
    """results.bin v1 image (include/emp_situation_binary_dumper.h:72-126)."""
    b = np.ascontiguousarray(np.asarray(boards, dtype=np.uint16).reshape(-1, width * height))
    ni = np.ascontiguousarray(np.asarray(node_index, dtype=np.uint64).reshape(-1))
    size = ctypes.c_uint64()
    _check(lib().emp_hip_results_bin_write(width, height, _ORDERS[order], 1 if solution_dump else 0, _ptr(b),
                                           _ptr(ni), b.shape[0], total_nodes, None, 0, ctypes.byref(size)))
    out = np.zeros(size.value, dtype=np.uint8)
    _check(lib().emp_hip_results_bin_write(width, height, _ORDERS[order], 1 if solution_dump else 0, _ptr(b),
                                           _ptr(ni), b.shape[0], total_nodes, _ptr(out), out.size,
                                           ctypes.byref(size)))
    return out.tobytes()


def results_bin_read(data):
    """Decodes a results.bin v1 image: dict with width, height,
    solution_dump, order [(x, y)], boards [n, P], node_index [n],
    total_nodes, complete."""
    buf = np.frombuffer(data, dtype=np.uint8).copy()
    w, h, sd, n, total, comp = (ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64(),
                                ctypes.c_uint64(), ctypes.c_uint32())
    _check(lib().emp_hip_results_bin_read(_ptr(buf), buf.size, ctypes.byref(w), ctypes.byref(h), ctypes.byref(sd),
                                          None, None, None, 0, ctypes.byref(n), ctypes.byref(total),
                                          ctypes.byref(comp)))
    P = w.value * h.value
    xy = np.zeros(2 * P, dtype=np.uint32)
    boards = np.zeros((n.value, P), dtype=np.uint16)
    ni = np.zeros(n.value, dtype=np.uint64)
    _check(lib().emp_hip_results_bin_read(_ptr(buf), buf.size, ctypes.byref(w), ctypes.byref(h), ctypes.byref(sd),
                                          _ptr(xy), _ptr(boards), _ptr(ni), n.value, ctypes.byref(n),
                                          ctypes.byref(total), ctypes.byref(comp)))
    return {"width": w.value, "height": h.value, "solution_dump": bool(sd.value),
            "order": [(int(xy[2 * i]), int(xy[2 * i + 1])) for i in range(P)], "boards": boards,
            "node_index": ni, "total_nodes": total.value, "complete": bool(comp.value)}


def order(width, height, kind="spiral"):
    xy = np.zeros(2 * width * height, dtype=np.uint32)
    _check(lib().emp_hip_order(width, height, _ORDERS[kind], _ptr(xy)))
    return [(int(xy[2 * i]), int(xy[2 * i + 1])) for i in range(width * height)]


def situation_string(width, height, board):
    board = np.ascontiguousarray(board, dtype=np.uint16)
    cap = width * height * 5 + 1
    buf = ctypes.create_string_buffer(cap)
    _check(lib().emp_hip_situation_string(width, height, _ptr(board), buf, cap))
    return buf.value.decode()


class Context:
    """One HIP stream on one device (emp_hip_context)."""

    def __init__(self, device=0):
        self._ctx = ctypes.c_void_p()
        _check(lib().emp_hip_create(device, ctypes.byref(self._ctx)))

    def close(self):
        if self._ctx:
            lib().emp_hip_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_info(self):
        cu, lds, clk = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib().emp_hip_device_info(self._ctx, ctypes.byref(cu), ctypes.byref(lds), ctypes.byref(clk)))
        return {"compute_units": cu.value, "lds_per_block": lds.value, "clock_khz": clk.value}

    def stream_copy(self, nbytes):
        """Device copy of nbytes with a plain kernel (counter calibration);
        returns the kernel time in ms."""
        ms = ctypes.c_double()
        _check(lib().emp_hip_stream_copy(self._ctx, nbytes, ctypes.byref(ms)))
        return ms.value

    def peaks(self):
        """Measured LDS read bandwidth (GB/s) and integer VALU throughput
        (G lane-ops/s) of the device."""
        lds, valu = ctypes.c_double(), ctypes.c_double()
        _check(lib().emp_hip_peaks(self._ctx, ctypes.byref(lds), ctypes.byref(valu)))
        return {"lds_gbs": lds.value, "valu_gops": valu.value}

    def mlp_train(self, features, targets, params, epochs=4, batch=1024, lr=1e-3, seed=1):
        """Adam on the scorer's MSE; returns (params, loss per epoch)."""
        x = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, MLP_INPUTS)
        y = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1)
        p = np.array(params, dtype=np.float32).reshape(MLP_PARAMS)
        losses = np.zeros(max(epochs, 1), dtype=np.float32)
        opt = MlpOptions(epochs, batch, lr, 0, seed)
        _check(lib().emp_hip_mlp_train(self._ctx, _ptr(x), _ptr(y), x.shape[0], ctypes.byref(opt), _ptr(p),
                                       _ptr(losses)))
        return p, losses[:epochs]

    def mlp_score(self, params, features):
        x = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, MLP_INPUTS)
        p = np.ascontiguousarray(params, dtype=np.float32).reshape(MLP_PARAMS)
        out = np.zeros(x.shape[0], dtype=np.float32)
        _check(lib().emp_hip_mlp_score(self._ctx, _ptr(p), _ptr(x), x.shape[0], _ptr(out)))
        return out

    def mlp_gradients(self, params, features, targets):
        """(loss, gradient of every parameter) over one batch of rows."""
        x = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, MLP_INPUTS)
        y = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1)
        p = np.ascontiguousarray(params, dtype=np.float32).reshape(MLP_PARAMS)
        g = np.zeros(MLP_PARAMS, dtype=np.float32)
        loss = ctypes.c_float()
        _check(lib().emp_hip_mlp_gradients(self._ctx, _ptr(p), _ptr(x), _ptr(y), x.shape[0], _ptr(g),
                                           ctypes.byref(loss)))
        return loss.value, g

    def glutton_max(self, width, height, colours, prefixes=None, push_budget=0, waves_per_cu=0,
                    n_searches=None, stacks=None, save_stacks=False):
        """One glutton_max search per prefix row (packed moves), or per
        saved stack (stacks = (state [n, state_words], masks [n, mask_words])
        from an earlier call); returns (results, boards [n, P] uint16, stats),
        and with save_stacks the stack of every search as it ended, as a
        fourth element."""
        colours = np.ascontiguousarray(colours, dtype=np.uint32)
        pz = Puzzle(width, height, colours.ctypes.data)
        P = width * height
        if stacks is not None:
            in_state = np.ascontiguousarray(stacks[0], dtype=np.uint32)
            in_masks = np.ascontiguousarray(stacks[1], dtype=np.uint64)
            n = in_state.shape[0]
            pre = np.zeros((n, 0), dtype=np.uint32)
        elif prefixes is None or len(prefixes) == 0 or np.asarray(prefixes).size == 0:
            n = n_searches or 1
            pre = np.zeros((n, 0), dtype=np.uint32)
        else:
            pre = np.ascontiguousarray(np.asarray(prefixes, dtype=np.uint32))
            n = pre.shape[0]
        out_state = out_masks = None
        if save_stacks:
            sw, mw = glutton_stack_words(P)
            out_state = np.zeros((n, sw), dtype=np.uint32)
            out_masks = np.zeros((n, mw), dtype=np.uint64)
        plen = pre.shape[1] if pre.ndim == 2 else 0
        res = (GluttonResult * n)()
        boards = np.zeros((n, width * height), dtype=np.uint16)
        opt = GluttonOptions(push_budget, waves_per_cu, 0)
        st = GluttonStats()
        _check(lib().emp_hip_glutton_max_stacks(
            self._ctx, ctypes.byref(pz), _ptr(pre) if plen else None, plen, n, ctypes.byref(opt),
            _ptr(in_state) if stacks is not None else None, _ptr(in_masks) if stacks is not None else None,
            _ptr(out_state) if save_stacks else None, _ptr(out_masks) if save_stacks else None,
            ctypes.byref(res), _ptr(boards), ctypes.byref(st)))
        results = [{"max_depth": r.max_depth, "status": r.status, "pushes": r.pushes} for r in res]
        stats = {f: getattr(st, f) for f, _ in GluttonStats._fields_}
        if save_stacks:
            return results, boards, stats, (out_state, out_masks)
        return results, boards, stats

    def glutton_wide(self, width, height, colours, parents=None, cells=None, capacity=1 << 16):
        """glutton_wide's breadth step: each parent (rows of packed moves from
        the level-0 capability; None: the level-0 situation alone) expanded
        at its cell (default: (0, 0) for every parent).  Returns a dict with
        parent_masks [n, P, nw] u64, parent_status, child_parent,
        child_move, child_masks [m, P, nw] and child_status."""
        colours = np.ascontiguousarray(colours, dtype=np.uint32)
        pz = Puzzle(width, height, colours.ctypes.data)
        P = width * height
        nw = (4 * P + 63) // 64
        if parents is None or np.asarray(parents).size == 0:
            n = 1 if parents is None else len(parents)
            rows = np.zeros((n, 0), dtype=np.uint32)
        else:
            rows = np.ascontiguousarray(np.asarray(parents, dtype=np.uint32))
            n = rows.shape[0]
        L = rows.shape[1]
        cl = np.ascontiguousarray(np.zeros(n, dtype=np.uint32) if cells is None else np.asarray(cells, dtype=np.uint32))
        pm = np.zeros((n, P, nw), dtype=np.uint64)
        ps = np.zeros(n, dtype=np.uint32)
        cp = np.zeros(capacity, dtype=np.uint32)
        cmv = np.zeros(capacity, dtype=np.uint32)
        cmask = np.zeros((capacity, P, nw), dtype=np.uint64)
        cs = np.zeros(capacity, dtype=np.uint32)
        nc = ctypes.c_uint64()
        _check(lib().emp_hip_glutton_wide(self._ctx, ctypes.byref(pz), _ptr(rows) if L else None, L, n, _ptr(cl),
                                          _ptr(pm), _ptr(ps), capacity, _ptr(cp), _ptr(cmv), _ptr(cmask), _ptr(cs),
                                          ctypes.byref(nc)))
        m = min(nc.value, capacity)
        return {"parent_masks": pm, "parent_status": ps, "n_children": nc.value, "child_parent": cp[:m].copy(),
                "child_move": cmv[:m].copy(), "child_masks": cmask[:m].copy(), "child_status": cs[:m].copy()}

    def situation_profile(self, width, height, colours, board):
        """Capability profiles of a situation (row-major board, 0xFFFF
        empty): level, glutton [10, P + 1] (total, max, min, vector_op,
        rvector_op; each min then max), surface [level + 1, 2P] sorted
        profiles, row_index [level + 1], n_variables."""
        colours = np.ascontiguousarray(colours, dtype=np.uint32)
        board = np.ascontiguousarray(board, dtype=np.uint16)
        pz = Puzzle(width, height, colours.ctypes.data)
        P = width * height
        level, nvar = ctypes.c_uint32(), ctypes.c_uint32()
        glutton = np.zeros((10, P + 1), dtype=np.uint32)
        surface = np.zeros((P + 1, 2 * P), dtype=np.uint32)
        row_index = np.zeros(P + 1, dtype=np.uint32)
        _check(lib().emp_hip_situation_profile(self._ctx, ctypes.byref(pz), _ptr(board), ctypes.byref(level),
                                               _ptr(glutton), _ptr(surface), _ptr(row_index), ctypes.byref(nvar)))
        L = level.value
        return {"level": L, "glutton": glutton, "surface": surface[:L + 1].copy(),
                "row_index": row_index[:L + 1].copy(), "n_variables": nvar.value}

    PROFILE_STATS = ("low_min", "low_max", "high_min", "high_max", "total_min", "total_max")

    def profile(self, width, height, colours, max_situations=1 << 24, solution_capacity=1 << 16):
        """The profile feature's search: stats [6, 3, P + 1] (PROFILE_STATS;
        series every situation / leading to a solution / the others; by
        level), the solutions (boards, ordered), level_sizes [P + 1] and the
        searches' starting situations."""
        colours = np.ascontiguousarray(colours, dtype=np.uint32)
        pz = Puzzle(width, height, colours.ctypes.data)
        P = width * height
        stats = np.zeros((6, 3, P + 1), dtype=np.uint32)
        sols = np.zeros((solution_capacity, P), dtype=np.uint16)
        sizes = np.zeros(P + 1, dtype=np.uint64)
        starts = np.zeros((6, P), dtype=np.uint16)
        n_sol, n_starts = ctypes.c_uint64(), ctypes.c_uint32()
        _check(lib().emp_hip_profile(self._ctx, ctypes.byref(pz), max_situations, _ptr(stats), _ptr(sols),
                                     solution_capacity, ctypes.byref(n_sol), _ptr(sizes), _ptr(starts),
                                     ctypes.byref(n_starts)))
        n = n_sol.value
        return {"stats": stats, "solutions": sols[:min(n, solution_capacity)].copy(), "n_solutions": n,
                "level_sizes": sizes, "starts": starts[:n_starts.value].copy()}

    def backtrack_trace(self, width, height, colours, order="basic", max_selections=0, trace_capacity=1 << 16):
        """The reference's CUDA backtracker: first solution and the selection
        trace as (lane, variable, step) rows."""
        colours = np.ascontiguousarray(colours, dtype=np.uint32)
        pz = Puzzle(width, height, colours.ctypes.data)
        board = np.zeros(width * height, dtype=np.uint16)
        trace = np.zeros((trace_capacity, 3), dtype=np.uint32)
        found, n = ctypes.c_uint32(), ctypes.c_uint64()
        _check(lib().emp_hip_backtrack_trace(self._ctx, ctypes.byref(pz), _ORDERS[order], max_selections,
                                             _ptr(board), ctypes.byref(found), _ptr(trace), trace_capacity,
                                             ctypes.byref(n)))
        return {"found": bool(found.value), "board": board, "trace": trace[:min(n.value, trace_capacity)].copy(),
                "trace_length": n.value}

    def expand(self, width, height, colours, depth, order="spiral", capacity=1 << 20):
        """Situations of the enumeration with `depth` pieces placed; returns
        (boards [n, P] uint16, n_situations, nodes_above, solutions_above)."""
        colours = np.ascontiguousarray(colours, dtype=np.uint32)
        pz = Puzzle(width, height, colours.ctypes.data)
        boards = np.zeros((capacity, width * height), dtype=np.uint16)
        n, nodes, sols = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().emp_hip_expand(self._ctx, ctypes.byref(pz), _ORDERS[order], depth, capacity, _ptr(boards),
                                    ctypes.byref(n), ctypes.byref(nodes), ctypes.byref(sols)))
        return boards[:min(n.value, capacity)].copy(), n.value, nodes.value, sols.value

    def _options(self, order, max_solutions, waves_per_cu, waves_per_block, steps_per_round, step_budget,
                 target_items, node_limit, queue_items, schedule="auto"):
        opt = EnumOptions()
        lib().emp_hip_enum_default_options(ctypes.byref(opt))
        opt.order = _ORDERS[order]
        opt.schedule = SCHEDULES[schedule]
        opt.waves_per_cu, opt.waves_per_block = waves_per_cu, waves_per_block
        opt.steps_per_round, opt.step_budget = steps_per_round, step_budget
        opt.target_items, opt.node_limit, opt.max_solutions = target_items, node_limit, max_solutions
        opt.queue_items = queue_items
        return opt

    def enumerate_items(self, width, height, colours, items=None, moves=None, order="spiral", max_solutions=1024,
                        ordered=False, time_limit_ms=0.0, rest_capacity=0, **launch):
        """Enumeration of the union of work items (items [n, 3] uint32
        floor/level/last, moves [n, P]; None: the whole enumeration).
        ordered: solutions in the sequential search's order with their node
        indices.  rest_capacity > 0 (with time_limit_ms or node_limit): the
        work left when a limit cut the run comes back as rest_items /
        rest_moves.
        Returns a dict with n_nodes, n_solutions, complete, max_depth, boards,
        node_index, rest_items, rest_moves and stats."""
        colours = np.ascontiguousarray(colours, dtype=np.uint32)
        pz = Puzzle(width, height, colours.ctypes.data)
        P = width * height
        # emp_hip_item is four u32: floor, level, last, reserved
        if items is None or len(items) == 0:
            n, it_arr, mv = 0, None, None
        else:
            items = np.asarray(items, dtype=np.uint32).reshape(-1, 3)
            mv = np.ascontiguousarray(np.asarray(moves, dtype=np.uint32).reshape(-1, P))
            if mv.shape[0] != items.shape[0]:
                raise ValueError("one row of moves per item")
            n = items.shape[0]
            it_arr = np.zeros((n, 4), dtype=np.uint32)
            it_arr[:, :3] = items
        opt = self._options(order, max_solutions, launch.get("waves_per_cu", 0), launch.get("waves_per_block", 0),
                            launch.get("steps_per_round", 0), launch.get("step_budget", 0),
                            launch.get("target_items", 0), launch.get("node_limit", 0), launch.get("queue_items", 0),
                            launch.get("schedule", "auto"))
        opt.time_limit_ms = time_limit_ms
        res = EnumResult()
        boards = np.zeros((max(max_solutions, 1), P), dtype=np.uint16)
        ni = np.zeros(max(max_solutions, 1), dtype=np.uint64)
        st = EnumStats()
        rest = np.zeros((rest_capacity, 4), dtype=np.uint32) if rest_capacity else None
        rest_mv = np.zeros((rest_capacity, P), dtype=np.uint32) if rest_capacity else None
        n_rest = ctypes.c_uint64()
        _check(lib().emp_hip_enumerate_items(self._ctx, ctypes.byref(pz), ctypes.byref(opt), _ptr(it_arr),
                                             _ptr(mv), n, ctypes.byref(res),
                                             _ptr(boards) if max_solutions else None,
                                             _ptr(ni) if ordered and max_solutions else None, _ptr(rest),
                                             _ptr(rest_mv), rest_capacity,
                                             ctypes.byref(n_rest) if rest_capacity else None, ctypes.byref(st)))
        out = {"n_nodes": res.n_nodes, "n_solutions": res.n_solutions, "complete": bool(res.complete),
               "max_depth": res.max_depth, "solve_ms": res.solve_ms, "boards": boards[:res.n_stored].copy(),
               "node_index": ni[:res.n_stored].copy() if ordered else None,
               "stats": {f: getattr(st, f) for f, _ in EnumStats._fields_}}
        if rest_capacity:
            m = n_rest.value
            out["rest_items"] = rest[:m, :3].copy()
            out["rest_moves"] = rest_mv[:m].copy()
        return out

    def probe(self, width, height, colours, rows, node_limit, order="spiral", time_limit_ms=0.0, **launch):
        """One search per row of moves (strategy indices 0..L-1) with
        node_limit placements each; returns (results, stats) with max_depth and
        n_nodes below each row."""
        colours = np.ascontiguousarray(colours, dtype=np.uint32)
        pz = Puzzle(width, height, colours.ctypes.data)
        rows = np.ascontiguousarray(np.asarray(rows, dtype=np.uint32))
        n, L = rows.shape
        opt = self._options(order, 0, launch.get("waves_per_cu", 0), launch.get("waves_per_block", 0),
                            launch.get("steps_per_round", 0), launch.get("step_budget", 0),
                            launch.get("target_items", 0), node_limit, launch.get("queue_items", 0),
                            launch.get("schedule", "auto"))
        opt.time_limit_ms = time_limit_ms
        res = (EnumResult * n)()
        st = EnumStats()
        _check(lib().emp_hip_probe(self._ctx, ctypes.byref(pz), _ptr(rows), L, n, ctypes.byref(opt),
                                   ctypes.byref(res), ctypes.byref(st)))
        out = [{"max_depth": r.max_depth, "n_nodes": r.n_nodes, "n_solutions": r.n_solutions,
                "complete": bool(r.complete)} for r in res]
        return out, {f: getattr(st, f) for f, _ in EnumStats._fields_}

    def enumerate(self, puzzles, order="spiral", max_solutions=1024, waves_per_cu=0,
                  waves_per_block=0, steps_per_round=0, step_budget=0, target_items=0,
                  node_limit=0, queue_items=0, ranks=None, time_limit_ms=0.0, use_graph=False,
                  noalloc=False, count_work=False, schedule="auto"):
        """Exhaustive enumeration of each (width, height, colours) puzzle.

        ranks (optional, one uint16 [P, T] array per puzzle) gives the visit
        order of the candidates at each strategy index.  Returns (results,
        stats): one dict per puzzle with n_nodes, n_solutions, complete,
        solve_ms, max_depth and boards ([n_stored, P] uint16 row-major).
        noalloc: through emp_hip_enumerate_noalloc (device memory from the
        library's fixed scratch buffer, no allocation).  schedule:
        "persistent" (one launch, hand-overs to idle lanes), "passes"
        (bounded launches, items split into the next one's queue) or "auto"
        (the first for one puzzle, the second for batches)."""
        n = len(puzzles)
        # The batch as flat arrays (one colour buffer, one board store) with
        # the C structs laid over numpy records: no per-puzzle ctypes objects.
        cols = [np.ascontiguousarray(c, dtype=np.uint32).ravel() for _, _, c in puzzles]
        sizes = np.array([w * h for w, h, _ in puzzles], dtype=np.int64)
        if (np.array([len(c) for c in cols], dtype=np.int64) != 4 * sizes).any():
            raise ValueError("each puzzle needs 4 colours per piece")
        flat = np.concatenate(cols) if n else np.zeros(1, dtype=np.uint32)
        arr = np.zeros(n, dtype=_PUZZLE_RECORD)
        arr["width"] = [w for w, _, _ in puzzles]
        arr["height"] = [h for _, h, _ in puzzles]
        arr["colours"] = flat.ctypes.data + 16 * (np.cumsum(sizes) - sizes)
        board_off = max_solutions * (np.cumsum(sizes) - sizes)
        store = np.zeros(max(1, int(max_solutions * sizes.sum())), dtype=np.uint16)
        board_ptrs = (store.ctypes.data + 2 * board_off).astype(np.uint64) if max_solutions else \
            np.zeros(n, dtype=np.uint64)
        opt = self._options(order, max_solutions, waves_per_cu, waves_per_block, steps_per_round, step_budget,
                            target_items, node_limit, queue_items, schedule)
        opt.time_limit_ms = time_limit_ms
        opt.use_graph = 1 if use_graph else 0
        opt.count_work = 1 if count_work else 0
        res = np.zeros(n, dtype=_ENUM_RESULT_RECORD)
        st = EnumStats()
        if noalloc:
            if ranks is not None:
                raise ValueError("the allocation-free enumeration takes no rank tables")
            _check(lib().emp_hip_enumerate_noalloc(self._ctx, arr.ctypes.data, n, ctypes.byref(opt),
                                                   res.ctypes.data, board_ptrs.ctypes.data, ctypes.byref(st)))
        elif ranks is None:
            _check(lib().emp_hip_enumerate(self._ctx, arr.ctypes.data, n, ctypes.byref(opt),
                                           res.ctypes.data, board_ptrs.ctypes.data, ctypes.byref(st)))
        else:
            rk = [np.ascontiguousarray(r, dtype=np.uint16) for r in ranks]
            if len(rk) != n:
                raise ValueError("one rank table per puzzle")
            for i, ((w, h, c), r) in enumerate(zip(puzzles, rk)):
                T = 4 * max(kind_counts(c))
                if r.shape != (w * h, T):
                    raise ValueError("rank table %d has shape %s, expected (%d, %d)" % (i, r.shape, w * h, T))
            rank_ptrs = (ctypes.c_void_p * n)(*[r.ctypes.data for r in rk])
            _check(lib().emp_hip_enumerate_ranked(self._ctx, arr.ctypes.data, n, ctypes.byref(opt), rank_ptrs,
                                                  res.ctypes.data, board_ptrs.ctypes.data, ctypes.byref(st)))
        stored = np.minimum(res["n_stored"], max_solutions).astype(np.int64)
        out = []
        for i, (nodes, sols, done, ms, depth, k, P, o) in enumerate(zip(
                res["n_nodes"].tolist(), res["n_solutions"].tolist(), res["complete"].tolist(),
                res["solve_ms"].tolist(), res["max_depth"].tolist(), stored.tolist(), sizes.tolist(),
                board_off.tolist())):
            out.append({"n_nodes": nodes, "n_solutions": sols, "complete": bool(done), "solve_ms": ms,
                        "max_depth": depth, "boards": store[o:o + k * P].reshape(k, P)})
        stats = {f: getattr(st, f) for f, _ in EnumStats._fields_}
        return out, stats
