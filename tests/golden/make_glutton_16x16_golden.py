"""Golden glutton_max results on the 16x16 shape (5 border / 17 interior
colours, generator seed 3), computed by the CPU oracle: two searches, each
from the planted centre hint plus 24 spiral moves, two pushes each.
Writes tests/golden/glutton_16x16_seed3.json.
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import edge_matching_puzzle_amd as emp  # noqa: E402
import oracle_bindings as o  # noqa: E402

W = H = 16
SEED, B, I = 3, 5, 17
BUDGET = 2


def main():
    c, planted = emp.generate(SEED, W, H, B, I)
    hint_pos = (H // 2) * W + (W // 2 - 1)
    hv = int(planted[hint_pos])
    hint = emp.pack_move(hint_pos, hv >> 2, hv & 3)
    rows = emp.spiral_prefixes(W, H, c, 24, 2, SEED, fixed=[hint])
    out = {"seed": SEED, "border_colours": B, "interior_colours": I, "push_budget": BUDGET,
           "searches": []}
    for row in rows:
        moves = [v for m in row for v in ((int(m) & 0xFF), (int(m) >> 8) & 0xFF, (int(m) >> 16) & 3)]
        r = o.glutton_max(W, H, c, prefix=moves, push_limit=BUDGET)
        out["searches"].append({"prefix": [int(m) for m in row], "max_depth": r["max_depth"],
                                "pushes": r["pushes"], "solved": r["solved"],
                                "budget_reached": r["status"] == o.LIMIT_REACHED,
                                "board": [int(v) for v in r["board"]]})
        print("search done", r["max_depth"], r["pushes"], flush=True)
    with open(os.path.join(HERE, "glutton_16x16_seed3.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
