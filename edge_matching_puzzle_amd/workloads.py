"""Synthetic workloads of the benchmark configurations (seeded, reproducible).

C0: 5x5, 3 border / 4 interior colours, one puzzle per seed.
C1: 8x8, 4 border / 8 interior colours, seed 7.
C2: 4096 puzzles, side uniform in 5..10, border colours in 3..5, interior
    colours in 4..12, one launch.
C3: 16x16, 5 border / 17 interior colours, seed 3; the planted piece at
    (7, 8) fixed as a hint, then prefixes of the first 24 spiral positions.
"""
import numpy as np

from . import generate, pack_move, spiral_prefixes

C2_MASTER_SEED = 20262


def c2_shapes(n=4096, master=C2_MASTER_SEED):
    """(side, border colours, interior colours, seed) of each C2 instance."""
    rng = np.random.default_rng(master)
    sides = rng.integers(5, 11, size=n)
    borders = rng.integers(3, 6, size=n)
    interiors = rng.integers(4, 13, size=n)
    return [(int(s), int(b), int(i), 2_000_000 + k) for k, (s, b, i) in enumerate(zip(sides, borders, interiors))]


def c2_puzzles(n=4096, master=C2_MASTER_SEED):
    out = []
    for side, b, i, seed in c2_shapes(n, master):
        c, _ = generate(seed, side, side, b, i)
        out.append((side, side, c))
    return out


def c3_puzzle(seed=3):
    w = h = 16
    c, planted = generate(seed, w, h, 5, 17)
    hint_pos = (h // 2) * w + (w // 2 - 1)
    hv = int(planted[hint_pos])
    return w, h, c, pack_move(hint_pos, hv >> 2, hv & 3)


def c3_prefixes(count, seed=3, length=24):
    w, h, c, hint = c3_puzzle(seed)
    return spiral_prefixes(w, h, c, length, count, seed, fixed=[hint])
