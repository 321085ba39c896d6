"""Capability profiles (the reference's situation_profile feature): the
oracle against the reference's own output files (tests/situation_profile_*/
references, copied under tests/golden/situation_profile), the HIP library
against the oracle, and the command line's files against the reference's
byte for byte."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle_bindings as o

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
PROFILES = os.path.join(GOLDEN, "situation_profile")
SITUATIONS = json.load(open(os.path.join(PROFILES, "situations.json")))
CRITERIA = ("total", "max", "min", "vector_op", "rvector_op")


def _board(w, h, text):
    """Row-major board of a situation string (ids right-aligned in cells of
    digits(P) + 1 characters, then N/E/S/W)."""
    P = w * h
    cw = len(str(P)) + 1
    assert len(text) == P * cw
    board = np.full(P, 0xFFFF, dtype=np.uint16)
    for c in range(P):
        cell = text[c * cw:(c + 1) * cw]
        if cell.strip("-") == "":
            continue
        board[c] = ((int(cell[:-1]) - 1) << 2) | "NESW".index(cell[-1])
    return board


def _rows(path, skip):
    lines = open(path).read().splitlines()[skip:]
    return [[int(v) for v in l.split()] for l in lines if l.strip()]


def _puzzle(name):
    return o.parse_puzzle(os.path.join(GOLDEN, "puzzles", name + ".txt"))


@pytest.mark.parametrize("name", sorted(SITUATIONS))
def test_oracle_matches_reference_profiles(name):
    w, h, c = _puzzle(name)
    P = w * h
    r = o.situation_profile(w, h, c, _board(w, h, SITUATIONS[name]))
    assert r["level"] == P
    d = os.path.join(PROFILES, name)
    # level 0: every variable counted once for its position, once for its piece
    assert r["surface"][0].sum() == 2 * r["n_variables"]
    assert _rows(os.path.join(d, name + "_level_glutton.txt"), 3) == r["glutton"].tolist()
    for i, crit in enumerate(CRITERIA):
        for j, mm in enumerate(("min", "max")):
            got = _rows(os.path.join(d, "%s_level_%s_%s_glutton.txt" % (name, crit, mm)), 3)
            assert got == [r["glutton"][2 * i + j].tolist()], (crit, mm)
    assert _rows(os.path.join(d, name + "_surface_vtk.txt"), 2) == r["surface"].tolist()
    offset = (P - r["row_index"].astype(np.int64))[:, None]
    assert _rows(os.path.join(d, name + "_line_plot_vtk.txt"), 3) == (r["surface"] + offset).tolist()


def test_oracle_profile_invariants():
    """On partial situations of generated puzzles: each placement empties its
    position and its piece (two more zeros per level), totals are twice the
    live variables and never grow, and the glutton totals start from the
    level-0 total."""
    import edge_matching_puzzle_amd as emp
    rng = np.random.default_rng(5)
    for seed, (w, h) in enumerate([(4, 4), (5, 5), (6, 4)]):
        c, sol = emp.generate(seed, w, h, 3, 4)
        P = w * h
        board = np.asarray(sol, dtype=np.uint16).copy()
        board[rng.permutation(P)[:P // 3]] = 0xFFFF
        r = o.situation_profile(w, h, c, board)
        L = r["level"]
        assert L == int((board != 0xFFFF).sum())
        tot = r["surface"].sum(axis=1)
        assert np.all(np.diff(tot.astype(np.int64)) <= 0)
        for k in range(L + 1):
            assert (r["surface"][k] == 0).sum() >= 2 * k
        assert np.all(r["glutton"][:, 0] == tot[0])
        assert np.all(r["glutton"][:, L + 1:] == 0)


def test_oracle_rejects_a_non_variable():
    w, h, c = _puzzle("3_3")
    board = _board(w, h, SITUATIONS["3_3"])
    v = int(board[0])
    board[0] = (v & ~3) | ((v + 1) & 3)   # the corner turned off the frame
    with pytest.raises(ValueError):
        o.situation_profile(w, h, c, board)


# generated puzzles whose solutions place only variables (a piece equal
# under a half turn keeps orientations 0 and 2 only, as in the reference)
GENERATED = [(4, 4, 3, 4, 1), (5, 5, 3, 4, 1), (6, 4, 3, 5, 2), (7, 5, 4, 5, 1), (6, 6, 3, 5, 1),
             (9, 7, 4, 6, 2), (8, 8, 4, 8, 2), (10, 10, 5, 9, 3)]


def _same(gpu, ref):
    assert gpu["level"] == ref["level"] and gpu["n_variables"] == ref["n_variables"]
    assert np.array_equal(gpu["glutton"], ref["glutton"])
    assert np.array_equal(gpu["surface"], ref["surface"])
    assert np.array_equal(gpu["row_index"], ref["row_index"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(SITUATIONS))
def test_gpu_matches_reference_situations(gpu_context, name):
    print("case: situation_profile", name, flush=True)
    w, h, c = _puzzle(name)
    b = _board(w, h, SITUATIONS[name])
    _same(gpu_context.situation_profile(w, h, c, b), o.situation_profile(w, h, c, b))


@pytest.mark.gpu
@pytest.mark.parametrize("w,h,B,I,seed", GENERATED)
def test_gpu_matches_oracle_on_generated(gpu_context, w, h, B, I, seed):
    """Full solutions and partial situations (a third, then two thirds of
    the cells emptied) of generated puzzles up to 10x10."""
    import edge_matching_puzzle_amd as emp
    c, sol = emp.generate(seed, w, h, B, I)
    P = w * h
    rng = np.random.default_rng(seed)
    full = np.asarray(sol, dtype=np.uint16)
    for drop in (0, P // 3, 2 * P // 3):
        b = full.copy()
        b[rng.permutation(P)[:drop]] = 0xFFFF
        print("case: situation_profile", w, h, seed, "empty cells", drop, flush=True)
        _same(gpu_context.situation_profile(w, h, c, b), o.situation_profile(w, h, c, b))


@pytest.mark.gpu
def test_gpu_16x16_profile_properties(gpu_context):
    """At 256 pieces (the reference's largest size) the oracle is too slow;
    the properties every profile has: level 0 counts every variable twice,
    each placement empties its position and piece, totals never grow, the
    full solution ends at 0, every glutton run starts at the level-0 total and
    its totals never grow while valid."""
    import edge_matching_puzzle_amd as emp
    w = h = 16
    c, sol = emp.generate(3, w, h, 5, 17)
    r = gpu_context.situation_profile(w, h, c, np.asarray(sol, dtype=np.uint16))
    assert r["level"] == 256
    s = r["surface"].astype(np.int64)
    assert s[0].sum() == 2 * r["n_variables"] == 2 * emp.backtrack_variables(w, h, c)
    assert np.all(np.diff(s.sum(axis=1)) <= 0) and s[-1].sum() == 0
    for k in range(0, 257, 16):
        assert (s[k] == 0).sum() >= 2 * k and np.all(np.diff(s[k]) >= 0)
    g = r["glutton"].astype(np.int64)
    assert np.all(g[:, 0] == 2 * r["n_variables"])
    for row in g:
        valid = row[1:] > 0
        head = row[:1 + int(np.argmin(valid)) if not valid.all() else 257]
        assert np.all(np.diff(head) <= 0)


@pytest.mark.gpu
def test_gpu_rejects_a_non_variable(gpu_context):
    import edge_matching_puzzle_amd as emp
    w, h, c = _puzzle("3_3")
    b = _board(w, h, SITUATIONS["3_3"])
    v = int(b[0])
    b[0] = (v & ~3) | ((v + 1) & 3)
    with pytest.raises(emp.EmpError):
        gpu_context.situation_profile(w, h, c, b)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(SITUATIONS))
def test_cli_writes_the_reference_files(tmp_path, name):
    """The command line of tests/situation_profile_<name>/test.info writes
    the reference's thirteen files byte for byte."""
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "edge_matching_puzzle_amd",
                       "edge_matching_puzzle")
    print("case: CLI situation_profile", name, flush=True)
    p = subprocess.run([exe, os.path.join(GOLDEN, "puzzles", name + ".txt"), "ressources", "situation_profile",
                        "--enable_gui=0", "--initial_situation=" + SITUATIONS[name]],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    ref = sorted(os.listdir(os.path.join(PROFILES, name)))
    assert sorted(os.listdir(tmp_path)) == ref
    for f in ref:
        assert open(os.path.join(tmp_path, f), "rb").read() == open(os.path.join(PROFILES, name, f), "rb").read(), f
    w, h, _ = _puzzle(name)
    lines = p.stdout.splitlines()
    assert str(w * h) in lines   # the situation's level
