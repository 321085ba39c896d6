"""emp_hip_enumerate_noalloc (device memory from the library's fixed scratch
buffer, no allocation) against emp_hip_enumerate, both through the C ABI.

Bar: bit-exact node counts, solution counts and sorted solution sets, on two
calls in a row (the second with a smaller input, so it runs on what the first
left in the buffer), and a clear error for a call beyond the buffer.
"""
import numpy as np
import pytest

import edge_matching_puzzle_amd as emp

pytestmark = pytest.mark.gpu


def _sorted(boards):
    return sorted(tuple(int(v) for v in b) for b in boards)


def _same(a, b):
    assert a["complete"] and b["complete"]
    assert (a["n_nodes"], a["n_solutions"], a["max_depth"]) == (b["n_nodes"], b["n_solutions"], b["max_depth"])
    assert _sorted(a["boards"]) == _sorted(b["boards"])


def _puzzles():
    big = []
    for seed in range(6):
        c, _ = emp.generate(700 + seed, 5, 5, 3, 4)
        big.append((5, 5, c))
    c, _ = emp.generate(3, 6, 6, 3, 6)
    big.append((6, 6, c))
    c, _ = emp.generate(11, 8, 8, 6, 16)
    big.append((8, 8, c))
    c, _ = emp.generate(5, 5, 5, 3, 4)
    small = [(5, 5, c)]
    return big, small


def test_noalloc_matches_enumerate_on_two_calls(gpu_context):
    big, small = _puzzles()
    print("case: noalloc, batch of %d then one 5x5" % len(big), flush=True)
    ref_big, _ = gpu_context.enumerate(big, max_solutions=512)
    ref_small, _ = gpu_context.enumerate(small, max_solutions=512)
    got_big, st = gpu_context.enumerate(big, max_solutions=512, noalloc=True)
    for a, b in zip(got_big, ref_big):
        _same(a, b)
    got_small, _ = gpu_context.enumerate(small, max_solutions=512, noalloc=True)
    _same(got_small[0], ref_small[0])


def test_noalloc_single_table_in_lds_and_forced_splits(gpu_context):
    """One puzzle (its table in LDS), a step budget small enough that items
    requeue many times: the queues in the scratch buffer hold them."""
    c, _ = emp.generate(1, 6, 6, 3, 6)
    print("case: noalloc 6x6 forced splits", flush=True)
    (ref,), _ = gpu_context.enumerate([(6, 6, c)], step_budget=64, target_items=1)
    (got,), st = gpu_context.enumerate([(6, 6, c)], step_budget=64, target_items=1, noalloc=True)
    assert st["splits"] > 100
    _same(got, ref)


def test_noalloc_refuses_a_call_beyond_the_buffer(gpu_context):
    c, _ = emp.generate(2, 5, 5, 3, 4)
    print("case: noalloc beyond the buffer", flush=True)
    with pytest.raises(emp.EmpError) as e:
        # 2^22 boards of 32 u16 are 256 MiB of solution store: more than the buffer
        gpu_context.enumerate([(5, 5, c)], max_solutions=1 << 22, noalloc=True)
    assert e.value.status == emp.OUT_OF_MEMORY
    assert "scratch" in str(e.value)
    # the context still works afterwards, with and without the buffer
    (a,), _ = gpu_context.enumerate([(5, 5, c)], noalloc=True)
    (b,), _ = gpu_context.enumerate([(5, 5, c)])
    _same(a, b)


def test_noalloc_refuses_graph_replay(gpu_context):
    c, _ = emp.generate(2, 5, 5, 3, 4)
    with pytest.raises(emp.EmpError) as e:
        gpu_context.enumerate([(5, 5, c)], use_graph=True, noalloc=True)
    assert e.value.status == emp.BAD_ARGUMENT
