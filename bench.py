"""Benchmark of the HIP search path on one MI355X.

Workload (one "step"): full exhaustive enumeration of one 8x8 synthetic
puzzle (64 pieces, 4 border colours, 8 interior colours, generator seed 7),
corner kind-index 0 fixed at (0,0), spiral order -- the plain backtracker
configuration the metric is quoted on.  The puzzle tables are built and
uploaded once before timing; each step runs the whole enumeration on the GPU.

Prints one JSON line:
  value            search nodes per second (placements, the reference's count)
  roofline         dominant kernel (dfs_kernel) against the LDS roofline: the
                   algorithmic LDS bytes per node (8 M (n + 3), M = ceil(4P/64)
                   u64 words per mask, n = placed neighbours of the next
                   position, averaged over the depth histogram) times nodes,
                   divided by the kernel time measured with HIP events on the
                   library's stream; peak = 256 CUs x 128 B/clk x 2.4 GHz.
  cpu_baseline     the CPU oracle (single thread) on the first nodes of the
                   same DFS.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "search nodes/s on 1 MI355X, batched+learned vs plain; max depth per GPU-minute"
LDS_PEAK_GBS = 256 * 128 * 2.4          # GB/s: 256 CUs x 128 B/clk x 2.4 GHz
VALU_PEAK_GOPS = 256 * 4 * 16 * 2.4     # 32-bit integer lane-ops/ns: 4 SIMD x 16 lanes per CU


def puzzle_c1(seed=7):
    import edge_matching_puzzle_amd as emp
    colours, _ = emp.generate(seed, 8, 8, 4, 8)
    return 8, 8, colours


def per_node_work(w, h, depth_nodes):
    """Algorithmic LDS bytes and integer ops per node, averaged over depth."""
    import edge_matching_puzzle_amd as emp
    order = emp.order(w, h)
    index = {xy: i for i, xy in enumerate(order)}
    P = w * h
    M = math.ceil(4 * P / 64)
    placed_nb = []
    for i, (x, y) in enumerate(order):
        n = 0
        for dx, dy in ((0, -1), (1, 0), (0, 1), (-1, 0)):
            nx, ny = x + dx, y + dy
            if 0 <= nx < w and 0 <= ny < h and index[(nx, ny)] < i:
                n += 1
        placed_nb.append(n)
    total = sum(depth_nodes)
    bytes_node = ops_node = 0.0
    for k, cnt in enumerate(depth_nodes):
        n_next = placed_nb[k + 1] if k + 1 < P else 0
        bytes_node += cnt * 8 * M * (n_next + 3)
        ops_node += cnt * (2 * M * (n_next + 1) + 2 * M)
    return bytes_node / total, ops_node / total, M


def cpu_baseline(w, h, colours, nodes):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bindings as o
    t = time.perf_counter()
    r = o.enumerate_(w, h, colours, node_limit=nodes, max_solutions=16)
    dt = time.perf_counter() - t
    return r, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--cpu-nodes", type=int, default=600_000_000,
                    help="nodes of the CPU baseline sample (about 15 s on one core)")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import edge_matching_puzzle_amd as emp
    w, h, colours = puzzle_c1(args.seed)
    ctx = emp.Context(0)
    for i in range(args.warmup):
        ctx.enumerate([(w, h, colours)], max_solutions=64)
        print("warmup step %d done" % i, file=sys.stderr, flush=True)
    nodes = 0
    kernel_ms = 0.0
    launches = 0
    results = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        (r,), st = ctx.enumerate([(w, h, colours)], max_solutions=64)
        nodes += r["n_nodes"]
        kernel_ms += st["dfs_kernel_ms"]
        launches += st["dfs_launches"]
        results.append((r, st))
        print("step done: %d nodes, %.1f ms kernel" % (r["n_nodes"], st["dfs_kernel_ms"]), file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    r, st = results[-1]
    if not r["complete"]:
        raise SystemExit("enumeration incomplete")
    value = nodes / elapsed

    cpu = None
    if not args.no_cpu:
        rc, dt = cpu_baseline(w, h, colours, args.cpu_nodes)
        depth = [int(v) for v in rc["depth_nodes"]]
        cpu = {"value": rc["n_nodes"] / dt, "unit": "nodes/s", "cores": 1, "kind": "port",
               "sample": "first %d nodes of the same DFS (CPU oracle, one thread, %.1f s)" % (rc["n_nodes"], dt)}
    else:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_bindings as o
        depth = [int(v) for v in o.enumerate_(w, h, colours, node_limit=20_000_000)["depth_nodes"]]
    bytes_node, ops_node, M = per_node_work(w, h, depth)
    achieved = nodes * bytes_node / (kernel_ms * 1e-3) / 1e9 if kernel_ms else 0.0
    valu = nodes * ops_node / (kernel_ms * 1e-3) / 1e9 if kernel_ms else 0.0
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "nodes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "C1 plain enumeration: one 8x8 puzzle, 4 border / 8 interior colours, "
                               "seed %d, corner kind-index 0 fixed, spiral order" % args.seed,
                   "nodes_per_step": r["n_nodes"], "solutions": r["n_solutions"],
                   "dfs_launches_per_step": st["dfs_launches"], "grid": [st["blocks"], st["threads"]],
                   "lds_tables": bool(st["lds_tables"])},
        "roofline": {"bound": "lds", "achieved": achieved, "peak": LDS_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / LDS_PEAK_GBS, "traffic": None,
                     "bytes_per_node": bytes_node, "words_per_mask": M,
                     "kernel_ms": kernel_ms, "kernel_launches": launches,
                     "valu_int_ops_per_node": ops_node, "valu_achieved_gops": valu,
                     "valu_peak_gops": VALU_PEAK_GOPS, "valu_frac": valu / VALU_PEAK_GOPS},
        "cpu_baseline": cpu,
    }
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
