"""Golden result of the C1 workload (one 8x8 puzzle, 4 border / 8 interior
colours, generator seed 7, spiral order) computed by the CPU oracle.

The oracle's DFS is cut at a depth into prefixes and each prefix's subtree is
enumerated in a process pool; the union of the subtrees is the whole
enumeration (checked against uncut runs in tests/test_oracle_golden.py).
Writes tests/golden/c1_8x8_seed7.json: node count, solution count and the
sorted solution strings.
"""
import json
import os
import sys
import time
from multiprocessing import Pool

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle_bindings as o  # noqa: E402

W = H = 8
SEED, B, I = 7, 4, 8
DEPTH = 10


def colours():
    import edge_matching_puzzle_amd as emp
    return emp.generate(SEED, W, H, B, I)[0]


def work(args):
    c, prefix = args
    r = o.subtree(W, H, c, prefix, max_solutions=256)
    return r["n_nodes"], r["n_solutions"], [o.situation_string(b, W * H) for b in r["boards"]]


def main():
    procs = int(sys.argv[1]) if len(sys.argv) > 1 else 7
    c = colours()
    t = time.time()
    pre, nodes, sols = o.prefixes(W, H, c, DEPTH)
    print("prefixes", len(pre), "nodes above cut", nodes, flush=True)
    boards = []
    done = 0
    with Pool(procs) as pool:
        for n, s, b in pool.imap_unordered(work, [(c, p) for p in pre], chunksize=8):
            nodes += n
            sols += s
            boards += b
            done += 1
            if done % 2000 == 0:
                print("%d/%d prefixes, %d nodes, %.0f s" % (done, len(pre), nodes, time.time() - t), flush=True)
    out = {"workload": "8x8, border colours 4, interior colours 8, seed 7, spiral order",
           "n_nodes": nodes, "n_solutions": sols, "solutions": sorted(boards),
           "cpu_seconds_wall": time.time() - t, "processes": procs}
    with open(os.path.join(HERE, "c1_8x8_seed7.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({k: v for k, v in out.items() if k != "solutions"}), flush=True)


if __name__ == "__main__":
    main()
