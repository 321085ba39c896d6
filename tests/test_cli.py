"""The C++ feature layer and command line (edge_matching_puzzle_amd/host)
against the reference's own CLI test expectations (tests/golden/expected_stdout.json,
collected from the reference's test.info files)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "edge_matching_puzzle_amd", "edge_matching_puzzle")
PUZZLES = os.path.join(ROOT, "tests", "golden", "puzzles")


def run(*args, timeout=120):
    p = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=timeout)
    return p.returncode, p.stdout


def test_unknown_feature_is_an_error():
    rc, out = run(os.path.join(PUZZLES, "3_3.txt"), "ressources", "no_such_feature")
    assert rc == 255 and out.splitlines()[0] == "9 pieces loaded"
    assert out.splitlines()[-1].startswith('ERROR : Unsupported feature "no_such_feature"')


@pytest.mark.parametrize("feature,flag", [("CUDA_glutton_max", "--dump_file=x.bin"),
                                          ("CUDA_backtracker", "--initial_situation=1N"),
                                          ("CUDA_glutton_wide", "--enable_gui=true"),
                                          ("new_strategy", "--strategy=strategy.txt"),
                                          ("new_strategy", "--no_such_flag=1")])
def test_flags_a_feature_does_not_use_are_errors(feature, flag):
    """Every flag of the reference's command line (src/main.cpp:51-70) is
    either used or refused before any device work: the GPU features take no
    dump file or initial situation, there is no GUI, only the basic and
    spiral orders exist."""
    rc, out = run(os.path.join(PUZZLES, "3_3.txt"), "ressources", feature, flag)
    assert rc == 255 and out.splitlines()[-1].startswith("ERROR : "), out


def test_refused_puzzle_is_an_error(tmp_path):
    bad = tmp_path / "bad.txt"
    text = open(os.path.join(PUZZLES, "3_3.txt")).read().replace("5,4,4,3,3", "5,4,4,3,9")
    bad.write_text(text)
    rc, out = run(str(bad), "ressources", "CUDA_glutton_max")
    assert rc == 255 and "multiple of 2" in out


def test_missing_piece_is_an_error(tmp_path):
    bad = tmp_path / "short.txt"
    bad.write_text("\n".join(open(os.path.join(PUZZLES, "3_3.txt")).read().splitlines()[:-1]) + "\n")
    rc, out = run(str(bad), "ressources", "CUDA_glutton_max")
    assert rc == 255 and "ERROR" in out


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["3_3", "4_4", "5_5"])
def test_glutton_max_feature_stdout(expected_stdout, name):
    """Every expected_stdout_string of tests/CUDA_glutton_max_<name>/test.info."""
    print("case: CLI CUDA_glutton_max", name, flush=True)
    rc, out = run(os.path.join(PUZZLES, name + ".txt"), "ressources", "CUDA_glutton_max", "--enable_gui=false")
    assert rc == 0, out
    for s in expected_stdout["CUDA_glutton_max_" + name]["strings"]:
        assert s in out


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["3_3", "4_4", "5_5"])
def test_backtracker_feature_stdout(expected_stdout, name):
    """Every expected_stdout_string of tests_CUDA/backtracker_<name>/test.info
    (the selection trace and the first solution)."""
    print("case: CLI CUDA_backtracker", name, flush=True)
    rc, out = run(os.path.join(PUZZLES, name + ".txt"), "ressources", "CUDA_backtracker", "--enable_gui=false")
    assert rc == 0, out
    lines = out.splitlines()
    for s in expected_stdout["backtracker_" + name]["strings"]:
        assert s in lines
    # prepare_transitions' report (include/system_equation_for_CUDA.h:81-82);
    # the counts of tests/system_equation_<name>/test.info:3
    nb = {"3_3": 36, "4_4": 144, "5_5": 484}[name]
    P = {"3_3": 9, "4_4": 16, "5_5": 25}[name]
    assert "Nb variables: %d" % nb in lines and "Max variables nb: %d" % (P * 1024) in lines


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["3_3", "4_4", "5_5"])
def test_enumerate_feature_totals(expected_stdout, name):
    """CUDA_enumerate lists the reference backtracker's solution among all
    solutions; totals equal the oracle's in the basic order."""
    import oracle_bindings as o
    print("case: CLI CUDA_enumerate", name, flush=True)
    rc, out = run(os.path.join(PUZZLES, name + ".txt"), "ressources", "CUDA_enumerate", "--enable_gui=false")
    assert rc == 0, out
    assert expected_stdout["backtracker_" + name]["strings"][-1] in out.splitlines()
    w, h, c = o.parse_puzzle(os.path.join(PUZZLES, name + ".txt"))
    ref = o.enumerate_(w, h, c, order=o.BASIC)
    assert "Total situations explored : %d" % ref["n_nodes"] in out
    assert "Nb solutions : %d" % ref["n_solutions"] in out


@pytest.mark.gpu
def test_glutton_wide_feature_levels():
    import oracle_bindings as o
    w, h, c = o.parse_puzzle(os.path.join(PUZZLES, "4_4.txt"))
    rc, out = run(os.path.join(PUZZLES, "4_4.txt"), "ressources", "CUDA_glutton_wide", "--hint=5")
    assert rc == 0, out
    for depth in range(2, 6):
        pre, _, _ = o.prefixes(w, h, c, depth)
        assert "Level %d : %d situations" % (depth, len(pre)) in out


def _kind_lists(c):
    kinds = {0: [], 1: [], 2: []}
    for p in range(len(c)):
        kinds[int(sum(1 for v in c[p] if v == 0))].append(p)
    return kinds


@pytest.mark.gpu
@pytest.mark.parametrize("w,B,I,seed,depth", [(5, 3, 4, 1, 6), (6, 3, 6, 2, 9), (8, 4, 8, 7, 10)])
def test_expand_matches_oracle_prefix_cut(gpu_context, w, B, I, seed, depth):
    """GPU breadth expansion = the oracle's DFS cut at the same depth."""
    import edge_matching_puzzle_amd as emp
    import oracle_bindings as o
    c, _ = emp.generate(seed, w, w, B, I)
    print("case: expand", w, depth, flush=True)
    boards, n, nodes, sols = gpu_context.expand(w, w, c, depth)
    pre, onodes, osols = o.prefixes(w, w, c, depth)
    assert n == len(pre) and nodes == onodes and sols == osols
    order = o.order(w, w)
    kinds = _kind_lists(c)
    ref = set()
    for row in pre:
        b = [0xFFFF] * (w * w)
        for k, t in enumerate(row):
            x, y = order[k]
            kind = 2 if (x in (0, w - 1) and y in (0, w - 1)) else (1 if (x in (0, w - 1) or y in (0, w - 1)) else 0)
            b[y * w + x] = (kinds[kind][int(t) >> 2] << 2) | (int(t) & 3)
        ref.add(tuple(b))
    assert {tuple(int(v) for v in b) for b in boards} == ref
