"""Host-side checks that need no GPU: the C ABI library loads and exports
every symbol include/emp_hip.h declares; generator, orders and formats."""
import os
import re

import numpy as np
import pytest

import edge_matching_puzzle_amd as emp
import oracle_bindings as o

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = []
    for f in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if f.endswith(".h"):
            text = open(os.path.join(ROOT, "include", f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"\b(emp_hip_\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = emp.lib()
    names = declared_symbols()
    assert len(names) >= 10
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(emp.SIGNATURES)


def test_version_string():
    assert b"gfx950" in emp.lib().emp_hip_version()


def test_generator_is_deterministic_and_seed_sensitive():
    a, pa = emp.generate(7, 8, 8, 4, 8)
    b, pb = emp.generate(7, 8, 8, 4, 8)
    c, _ = emp.generate(8, 8, 8, 4, 8)
    assert np.array_equal(a, b) and np.array_equal(pa, pb)
    assert not np.array_equal(a, c)


def test_generator_colour_ranges():
    c, _ = emp.generate(3, 10, 10, 5, 12)
    frame_side = c.min(axis=1) == 0
    assert c.max() <= 17
    # pieces with a frame side carry border colours on their frame-adjacent sides
    assert frame_side.sum() == 36


@pytest.mark.parametrize("w,h", [(3, 3), (4, 4), (5, 5), (8, 8), (16, 16), (3, 4), (6, 4), (10, 7), (4, 6)])
@pytest.mark.parametrize("kind", ["spiral", "basic"])
def test_orders_match_oracle(w, h, kind):
    ref = o.order(w, h, o.SPIRAL if kind == "spiral" else o.BASIC)
    if len(set(ref)) == w * h:
        assert emp.order(w, h, kind) == ref
    else:
        # include/emp_spiral_strategy_generator.h:52-98 revisits cells on
        # boards with height >= width + 2; the library refuses that order.
        assert kind == "spiral" and h >= w + 2
        with pytest.raises(emp.EmpError) as e:
            emp.order(w, h, kind)
        assert e.value.status == emp.UNSUPPORTED


def test_oracle_refuses_non_permutation_spiral():
    c, _ = emp.generate(1, 4, 6, 2, 3)
    with pytest.raises(ValueError):
        o.enumerate_(4, 6, c)
    assert o.enumerate_(4, 6, c, order=o.BASIC)["n_nodes"] > 0


def test_situation_string_matches_reference_format(golden_dir):
    w, h, c = emp.parse_puzzle(os.path.join(golden_dir, "puzzles", "5_5.txt"))
    r = o.enumerate_(w, h, c)
    assert emp.situation_string(w, h, r["boards"][0]) == \
        " 1N15E 3N10E21S20S18W17E 7S 4W 2E 9E 8S14N 6S23W13E19E12W16S 5E22N11E24N25N"
    empty = np.full(9, 0xFFFF, dtype=np.uint16)
    assert emp.situation_string(3, 3, empty) == "-" * 18


def test_parse_puzzle_matches_oracle_parser(golden_dir):
    for n in ("3_3", "4_4", "5_5"):
        p = os.path.join(golden_dir, "puzzles", n + ".txt")
        a, b = emp.parse_puzzle(p), o.parse_puzzle(p)
        assert a[0] == b[0] and a[1] == b[1] and np.array_equal(a[2], b[2])


def test_puzzle_validation(golden_dir):
    """Refusals of include/emp_piece_db.h:898-924 and kind counts."""
    w, h, c = emp.parse_puzzle(os.path.join(golden_dir, "puzzles", "3_3.txt"))
    emp.validate(w, h, c)
    bad = c.copy()
    bad[4, 0] = 9                      # colour 9 on one edge only
    with pytest.raises(emp.EmpError) as e:
        emp.validate(w, h, bad)
    assert e.value.status == emp.BAD_PUZZLE and "multiple of 2" in str(e.value)
    bad = c.copy()
    bad[4] = [0, 0, 0, 1]              # three frame sides
    with pytest.raises(emp.EmpError):
        emp.validate(w, h, bad)
    with pytest.raises(emp.EmpError):
        emp.validate(4, 4, np.zeros((16, 4), np.uint32))


def test_null_context_is_an_argument_error():
    with pytest.raises(emp.EmpError) as e:
        emp._check(emp.lib().emp_hip_enumerate(None, None, 0, None, None, None, None))
    assert e.value.status == emp.BAD_ARGUMENT


def test_spiral_prefixes_are_valid_and_distinct():
    w = h = 8
    c, planted = emp.generate(5, w, h, 4, 8)
    hint = emp.pack_move(36, int(planted[36]) >> 2, int(planted[36]) & 3)
    rows = emp.spiral_prefixes(w, h, c, 10, 50, 9, fixed=[hint])
    assert len(rows) == 50 and len({tuple(r) for r in rows}) == 50
    order = emp.order(w, h)
    for row in rows:
        assert row[0] == hint
        cells = {}
        for k, m in enumerate(row[1:]):
            pos, piece, ori = int(m) & 0xFF, (int(m) >> 8) & 0xFF, (int(m) >> 16) & 3
            assert pos == order[k][1] * w + order[k][0]
            cells[pos] = [int(c[piece][(s + ori) % 4]) for s in range(4)]
        pieces = [(int(m) >> 8) & 0xFF for m in row]
        assert len(set(pieces)) == len(pieces)
        for pos, look in cells.items():
            x, y = pos % w, pos // w
            assert (look[0] == 0) == (y == 0) and (look[3] == 0) == (x == 0)
            assert (look[1] == 0) == (x == w - 1) and (look[2] == 0) == (y == h - 1)
            if pos + 1 in cells and x + 1 < w:
                assert look[1] == cells[pos + 1][3]
            if pos + w in cells:
                assert look[2] == cells[pos + w][0]


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with a message before
    touching a device."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr + r.stdout
    assert r.stdout.strip() == ""


@pytest.mark.parametrize("name,count", [("3_3", 36), ("4_4", 144), ("5_5", 484)])
def test_backtracker_variable_counts(golden_dir, name, count):
    """tests/system_equation_<name>/test.info:3 -- "Nb variables : 36 / 144 /
    484" -- counted from the backtracker's level-0 capability."""
    w, h, c = emp.parse_puzzle(os.path.join(golden_dir, "puzzles", name + ".txt"))
    assert emp.backtrack_variables(w, h, c) == count
    assert emp.backtrack_variables(w, h, c, order="spiral") == count


def test_variable_count_of_synthetic_boards():
    """Counted independently: each corner cell takes each corner piece in one
    orientation, each border cell each border piece in one, each centre cell
    each centre piece in its kept orientations (4, 2 when N = S and E = W,
    1 when all four sides match; include/emp_piece_db.h:1266-1274)."""
    for w, h, B, I, seed in ((5, 5, 3, 4, 1), (8, 8, 4, 8, 7), (7, 5, 3, 5, 2), (6, 6, 1, 2, 3)):
        c, _ = emp.generate(seed, w, h, B, I)
        zeros = (c == 0).sum(axis=1)
        nb, nc = int((zeros == 1).sum()), int((zeros == 0).sum())
        kept = 0
        for p in np.nonzero(zeros == 0)[0]:
            n, e, s_, wst = (int(v) for v in c[p])
            kept += 1 if n == e == s_ == wst else (2 if (n == s_ and e == wst) else 4)
        assert emp.backtrack_variables(w, h, c) == 16 + nb * nb + nc * kept
