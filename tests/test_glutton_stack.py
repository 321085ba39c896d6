"""Glutton stack checkpoints: a search cut at its push budget saves its stack
(the reference's CUDA_glutton_stack_XML_converter layout), and the search
resumed from that stack continues exactly as the uncut search does -- the
same moves, max depth, status and board, with the pushes adding up.  The XML
form round-trips, and malformed or inconsistent stacks are refused with the
reference loader's messages before any device work."""
import json
import os

import numpy as np
import pytest

import edge_matching_puzzle_amd as emp
import oracle_bindings as o
from board_check import board_problems

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _level0_stack(w, h, c):
    """The stack of a search that has placed nothing: identity slot maps,
    every piece available, the level-0 capability as level 0's masks."""
    P = w * h
    sw, mw = emp.glutton_stack_words(P)
    nw = (4 * P + 63) // 64
    st = np.zeros(sw, dtype=np.uint32)
    st[0] = st[1] = P
    st[5:5 + P] = np.arange(P)                   # info_to_position
    st[5 + P:5 + 2 * P] = np.arange(P)           # position_to_info
    st[5 + 2 * P:5 + 4 * P] = 0xFFFFFFFF         # played, max_played
    for q in range(P):
        st[5 + 4 * P + q // 32] |= np.uint32(1 << (q % 32))
    masks0, status = o.glutton_masks(w, h, c)
    assert status == 0
    mk = np.zeros(mw, dtype=np.uint64)
    mk[:P * nw] = masks0.reshape(-1)
    return st, mk


def test_stack_sizes():
    for P in (9, 64, 256):
        sw, mw = emp.glutton_stack_words(P)
        assert sw == 13 + 4 * P
        assert mw == P * (P + 1) // 2 * ((4 * P + 63) // 64)
    with pytest.raises(emp.EmpError):
        emp.glutton_stack_words(257)


def test_xml_round_trip_and_layout():
    w, h, c = o.parse_puzzle(os.path.join(GOLDEN, "puzzles", "4_4.txt"))
    st, mk = _level0_stack(w, h, c)
    xml = emp.glutton_stack_to_xml(w, h, c, st, mk)
    assert xml.startswith('<CUDA_glutton_stack size="16" level="0" nb_pieces="16" max="0" base="0">')
    for name, size in (("info_to_position", 16), ("position_to_info", 16), ("played_info", 16),
                       ("max_played_info", 0), ("available_pieces", 8)):
        assert '<%s size="%d">' % (name, size) in xml
    assert xml.count("<level index=") == 16 and xml.count("<position_info index=") == 16 * 17 // 2
    assert xml.count("<word index=") == 32 * 16 * 17 // 2
    assert '<item index="0">0x0000ffff</item>' in xml   # 16 pieces available
    st2, mk2 = emp.glutton_stack_from_xml(w, h, c, xml)
    assert np.array_equal(st, st2) and np.array_equal(mk, mk2)
    # declarations, comments, other white space and attribute order are read too
    alt = '<?xml version="1.0"?>\n<!-- saved -->\n' + xml.replace("\n", "\r\n  ").replace(
        'size="16" level="0"', 'level="0"   size="16"')
    st3, mk3 = emp.glutton_stack_from_xml(w, h, c, alt)
    assert np.array_equal(st, st3) and np.array_equal(mk, mk3)
    # a file without the base attribute (the reference's) resumes above level 0
    st4, _ = emp.glutton_stack_from_xml(w, h, c, xml.replace(' base="0"', ""))
    assert st4[4] == 0


@pytest.mark.parametrize("edit,message", [
    (lambda x: x.replace('size="16" level', 'size="15" level', 1), "size from XML (15) is different from size of stack (16)"),
    (lambda x: x.replace(' max="0"', ""), "Missing max attribute in CUDA_glutton_stack node"),
    (lambda x: x[:x.index("<played_info")] + x[x.index("</played_info>") + len("</played_info>"):],
     'No node "played_info" in node "CUDA_glutton_stack"'),
    (lambda x: x.replace('<position_to_info size="16">\n    <item index="0">0x00000000</item>',
                         '<position_to_info size="16">\n    <item index="0">0x00000003</item>'),
     "disagree"),
    (lambda x: x.replace('<item index="1">0x00000001</item>', '<item index="2">0x00000001</item>', 1),
     "index from XML (2) is different from index of info_to_position item (1)"),
    (lambda x: x.replace('<word index="2">0x00000000</word>', '<word index="2">0x00000001</word>', 1),
     "bit past the last piece"),
    (lambda x: x.replace("</CUDA_glutton_stack>", ""), "is not closed"),
    (lambda x: x.replace('<item index="0">0x0000ffff</item>', '<item index="0">0x0001ffff</item>'),
     "available piece past the last piece"),
])
def test_xml_errors(edit, message):
    w, h, c = o.parse_puzzle(os.path.join(GOLDEN, "puzzles", "4_4.txt"))
    st, mk = _level0_stack(w, h, c)
    bad = edit(emp.glutton_stack_to_xml(w, h, c, st, mk))
    with pytest.raises(emp.EmpError) as e:
        emp.glutton_stack_from_xml(w, h, c, bad)
    assert message in str(e.value), str(e.value)


def _golden_16x16():
    return json.load(open(os.path.join(GOLDEN, "glutton_16x16_seed3.json")))


def _same(a, b, ba, bb):
    assert (a["status"], a["max_depth"]) == (b["status"], b["max_depth"])
    assert np.array_equal(ba, bb)


def _cut_and_resume(ctx, w, h, c, rows, cuts):
    """The searches of `rows` run uncut for sum(cuts) pushes, and cut into
    len(cuts) calls, each resumed from the stacks the previous call saved
    (through the XML form for one of them)."""
    total = sum(cuts)
    ref, ref_boards, _ = ctx.glutton_max(w, h, c, prefixes=rows, push_budget=total)
    stacks = None
    pushes = np.zeros(len(rows), dtype=np.int64)
    for k, b in enumerate(cuts):
        if stacks is None:
            res, boards, _, stacks = ctx.glutton_max(w, h, c, prefixes=rows, push_budget=b, save_stacks=True)
        else:
            if k == 1:
                # the second resume goes through the XML form of every stack
                xs = [emp.glutton_stack_from_xml(w, h, c, emp.glutton_stack_to_xml(w, h, c, s, m))
                      for s, m in zip(*stacks)]
                assert all(np.array_equal(x[0], s) and np.array_equal(x[1], m) for x, s, m in zip(xs, *stacks))
            res, boards, _, stacks = ctx.glutton_max(w, h, c, push_budget=b, stacks=stacks, save_stacks=True)
        pushes += np.array([r["pushes"] for r in res])
    for i in range(len(rows)):
        _same(res[i], ref[i], boards[i], ref_boards[i])
        assert pushes[i] == ref[i]["pushes"], i
        assert not board_problems(w, h, c, boards[i], prefix_moves=rows[i], depth=res[i]["max_depth"],
                                  solved=res[i]["status"] == emp.GLUTTON_SOLVED)
    return ref


@pytest.mark.gpu
def test_cut_and_resume_8x8(gpu_context):
    """Searches from short spiral prefixes of 8x8 puzzles, cut after 3, 5 and
    then 200 pushes: the same results as one uncut run of 208 pushes, which
    is the oracle's."""
    for seed in (1, 2):
        w = h = 8
        c, _ = emp.generate(seed, w, h, 4, 6)
        rows = emp.spiral_prefixes(w, h, c, 6, 4, seed)
        print("case: glutton cut and resume 8x8 seed", seed, flush=True)
        ref = _cut_and_resume(gpu_context, w, h, c, rows, [3, 5, 200])
        for i, row in enumerate(rows[:2]):
            moves = [v for m in row for v in ((int(m) & 0xFF), (int(m) >> 8) & 0xFF, (int(m) >> 16) & 3)]
            r = o.glutton_max(w, h, c, prefix=moves, push_limit=208)
            assert (r["max_depth"], r["pushes"]) == (ref[i]["max_depth"], ref[i]["pushes"])


@pytest.mark.gpu
def test_cut_and_resume_16x16(gpu_context):
    """The 16x16 golden's start situations (5 border / 17 interior colours,
    centre hint + 24 spiral moves), cut after 4, 5 and 7 pushes: the golden's
    16-push results."""
    g = _golden_16x16()
    w = h = 16
    c, _ = emp.generate(g["seed"], w, h, g["border_colours"], g["interior_colours"])
    rows = np.array([s["prefix"] for s in g["searches"][:6]], dtype=np.uint32)
    print("case: glutton cut and resume 16x16", flush=True)
    ref = _cut_and_resume(gpu_context, w, h, c, rows, [4, 5, 7])
    for r, s in zip(ref, g["searches"][:6]):
        assert (r["max_depth"], r["pushes"]) == (s["max_depth"], s["pushes"])


@pytest.mark.gpu
def test_resume_of_a_finished_search(gpu_context):
    """A solved or exhausted search's stack resumes to the same result with
    no push."""
    w, h, c = o.parse_puzzle(os.path.join(GOLDEN, "puzzles", "5_5.txt"))
    (r,), b, _, stacks = gpu_context.glutton_max(w, h, c, save_stacks=True)
    assert r["status"] == emp.GLUTTON_SOLVED
    (r2,), b2, _ = gpu_context.glutton_max(w, h, c, stacks=stacks)
    assert r2["status"] == emp.GLUTTON_SOLVED and r2["pushes"] == 0 and np.array_equal(b, b2)


@pytest.mark.gpu
def test_inconsistent_stack_is_refused(gpu_context):
    w, h, c = o.parse_puzzle(os.path.join(GOLDEN, "puzzles", "4_4.txt"))
    st, mk = _level0_stack(w, h, c)
    st[2] = 3   # level 3 with nothing placed
    with pytest.raises(emp.EmpError):
        gpu_context.glutton_max(w, h, c, stacks=(st[None, :], mk[None, :]))
    st, mk = _level0_stack(w, h, c)
    (r,), _, _ = gpu_context.glutton_max(w, h, c, stacks=(st[None, :], mk[None, :]))
    (r0,), _, _ = gpu_context.glutton_max(w, h, c)
    assert (r["status"], r["max_depth"], r["pushes"]) == (r0["status"], r0["max_depth"], r0["pushes"])
