"""Golden glutton_max results on the 16x16 shape (5 border / 17 interior
colours, generator seed 3), computed by the CPU oracle: the first 16 of the
benchmark's start situations (the planted centre hint plus 24 spiral moves,
workloads.c3_prefixes), 16 pushes each.  The oracle needs 20-40 s per push at
this size, so the searches run in a pool of worker processes.
Writes tests/golden/glutton_16x16_seed3.json.
"""
import json
import os
import sys
from multiprocessing import Pool

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

SEED, B, I = 3, 5, 17
N_SEARCHES = 16
BUDGET = 16
WORKERS = 6


def _one(row):
    import oracle_bindings as o
    from edge_matching_puzzle_amd.workloads import c3_puzzle
    w, h, c, _ = c3_puzzle(SEED)
    moves = [v for m in row for v in ((int(m) & 0xFF), (int(m) >> 8) & 0xFF, (int(m) >> 16) & 3)]
    r = o.glutton_max(w, h, c, prefix=moves, push_limit=BUDGET)
    print("search done: depth %d, %d pushes" % (r["max_depth"], r["pushes"]), flush=True)
    return {"prefix": [int(m) for m in row], "max_depth": int(r["max_depth"]), "pushes": int(r["pushes"]),
            "solved": bool(r["solved"]), "budget_reached": r["status"] == o.LIMIT_REACHED,
            "board": [int(v) for v in r["board"]]}


def main():
    from edge_matching_puzzle_amd.workloads import c3_prefixes
    rows = c3_prefixes(N_SEARCHES, SEED)
    with Pool(WORKERS) as pool:
        searches = pool.map(_one, [r.tolist() for r in rows])
    out = {"seed": SEED, "border_colours": B, "interior_colours": I, "push_budget": BUDGET,
           "searches": searches}
    with open(os.path.join(HERE, "glutton_16x16_seed3.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
