"""Writes c4_exact_8x8.json: oracle node and solution counts, and a hash of
the sorted solution boards, for the 256 8x8 instances (8 border / 10
interior colours) the learned-ordering exactness check enumerates.

Run from the repository root: python tests/golden/make_c4_exact_golden.py
"""
import hashlib
import json
import os
import sys
from multiprocessing import Pool

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def board_hash(boards):
    import numpy as np
    b = np.asarray(boards, dtype=np.uint16).reshape(len(boards), -1)
    rows = sorted(bytes(r.tobytes()) for r in b)
    return hashlib.sha256(b"".join(rows)).hexdigest()


def one(seed):
    import oracle_bindings as o
    from edge_matching_puzzle_amd import learned
    w, h, c = learned.puzzle(seed, **learned.EXACT_SHAPE)
    r = o.enumerate_(w, h, c, max_solutions=4096)
    assert r["n_solutions"] <= 4096
    return {"seed": seed, "n_nodes": int(r["n_nodes"]), "n_solutions": int(r["n_solutions"]),
            "boards_sha256": board_hash(r["boards"]) if r["n_solutions"] else ""}


def main():
    from edge_matching_puzzle_amd import learned
    with Pool(min(7, os.cpu_count() or 1)) as pool:
        rows = pool.map(one, learned.EXACT_SEEDS)
    out = {"shape": learned.EXACT_SHAPE, "order": "spiral", "instances": rows,
           "total_nodes": sum(r["n_nodes"] for r in rows)}
    path = os.path.join(ROOT, "tests", "golden", "c4_exact_8x8.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path, "total nodes", out["total_nodes"])


if __name__ == "__main__":
    main()
