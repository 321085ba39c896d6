"""Benchmark of the HIP search path on one MI355X.

Default line (the metric's plain leg, with the batched leg beside it):
  one step = full exhaustive enumeration of one 8x8 synthetic puzzle (64
  pieces, 4 border colours, 8 interior colours, generator seed 7), corner
  kind-index 0 fixed at (0,0), spiral order -- the plain backtracker
  configuration the metric is quoted on.  Each step runs the whole
  enumeration on the GPU (tables built and uploaded by the call, breadth
  expansions, then one persistent DFS launch).
  batched: one step = 4096 synthetic puzzles of side 5..10, 3..5 border and
  4..12 interior colours in one enumeration (tables built on host threads,
  DFS passes replayed from hipGraphs, items regrouped by puzzle before each
  pass and taken by each XCD from its own region in per-wave runs), each stopped near
  2e7 nodes (the cells with more than 10^9 expected solutions cannot be
  enumerated by anyone; the count of complete and cut instances is
  reported).

Prints one JSON line:
  value            search nodes per second (placements, the reference's count)
  roofline         dominant kernel (dfs_kernel) against the LDS roofline: the
                   algorithmic LDS bytes of the enumeration, 8 M (n + 3) per
                   placement (M = ceil(4P/64) u64 words per mask, n = placed
                   neighbours of the next position), counted exactly by an
                   instrumented run of the same enumeration outside the timed
                   region, per node times the timed nodes, divided by the
                   kernel time measured with HIP events on the library's
                   stream; peak = 256 CUs x 256 B/clk x 2.4 GHz.
                   peak_measured / valu_peak_measured_gops: the same peaks
                   measured on the device by stream kernels (emp_hip_peaks).
  cpu_baseline     the CPU oracle (single thread) on the first nodes of the
                   same DFS.
  batched          the same fields for the batched leg (cpu_baseline: the
                   oracle in a pool of host processes, one instance each);
                   its launch is picked by the autotuner on the batch itself
                   before timing (tuned, autotune_trials; --no-autotune: the
                   default launch).
  learned          the learned-ordering leg at a reduced size: a scorer
                   trained on 8 seeds, then the mean max depth on 4 held-out
                   16x16 seeds in equal device time (300 ms) for the plain,
                   trained and random orders.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "search nodes/s on 1 MI355X, batched+learned vs plain; max depth per GPU-minute"
LDS_PEAK_GBS = 256 * 256 * 2.4          # GB/s: 256 CUs x 256 B/clk (CDNA4 LDS reads) x 2.4 GHz
VALU_PEAK_GOPS = 256 * 4 * 16 * 2.4     # 32-bit integer lane-ops/ns: 4 SIMD x 16 lanes per CU
C2_NODE_LIMIT = 20_000_000              # per instance, batched leg


def puzzle_c1(seed=7):
    import edge_matching_puzzle_amd as emp
    colours, _ = emp.generate(seed, 8, 8, 4, 8)
    return 8, 8, colours


def cpu_baseline(w, h, colours, nodes):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bindings as o
    t = time.perf_counter()
    r = o.enumerate_(w, h, colours, node_limit=nodes, max_solutions=16)
    dt = time.perf_counter() - t
    return r, dt


def _cpu_pool_c2(puzzles, limit, seconds=15.0):
    """Oracle on the host cores, one instance per worker, bounded sample."""
    from multiprocessing import Pool
    cores = min(16, os.cpu_count() or 1)
    t = time.perf_counter()
    nodes = 0
    done = 0
    with Pool(cores) as pool:
        for r in pool.imap_unordered(_cpu_one, [(w, h, c.tolist(), limit) for w, h, c in puzzles]):
            nodes += r
            done += 1
            if time.perf_counter() - t > seconds:
                break
        pool.terminate()
    dt = time.perf_counter() - t
    return {"value": nodes / dt, "unit": "nodes/s", "cores": cores, "kind": "port",
            "sample": "first %d of the batch's instances, node limit %d each, one instance per process (%.1f s)"
                      % (done, limit, dt)}


def _cpu_one(args):
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bindings as o
    w, h, c, limit = args
    return o.enumerate_(w, h, np.array(c, dtype=np.uint32), node_limit=limit, max_solutions=1)["n_nodes"]


def _roofline(nodes, kernel_ms, work, launches, peaks=None):
    """LDS roofline of dfs_kernel: work = (nodes, bytes, ops) of an
    instrumented run of the same enumeration."""
    w_nodes, w_bytes, w_ops = work
    bytes_node, ops_node = w_bytes / w_nodes, w_ops / w_nodes
    achieved = nodes * bytes_node / (kernel_ms * 1e-3) / 1e9 if kernel_ms else 0.0
    valu = nodes * ops_node / (kernel_ms * 1e-3) / 1e9 if kernel_ms else 0.0
    r = {"bound": "lds", "achieved": achieved, "peak": LDS_PEAK_GBS, "unit": "GB/s",
         "frac": achieved / LDS_PEAK_GBS, "traffic": None, "bytes_per_node": bytes_node,
         "kernel_ms": kernel_ms, "kernel_launches": launches, "valu_int_ops_per_node": ops_node,
         "valu_achieved_gops": valu, "valu_peak_gops": VALU_PEAK_GOPS, "valu_frac": valu / VALU_PEAK_GOPS}
    if peaks:
        r.update({"peak_measured": peaks["lds_gbs"], "frac_of_measured": achieved / peaks["lds_gbs"],
                  "valu_peak_measured_gops": peaks["valu_gops"], "valu_frac_of_measured": valu / peaks["valu_gops"]})
    return r


# Launch space the batched leg is tuned over (the middle of each axis is the
# library's default launch, where the coordinate search starts).
BATCHED_SPACE = {
    "waves_per_cu": [12, 16],
    "waves_per_block": [2, 4],
    "steps_per_round": [64, 128, 256],
    "target_per_lane": [2, 4, 8],
    "step_budget": [1 << 14, 1 << 16, 1 << 18],
}


def bench_batched(ctx, steps, warmup, cpu=None, instances=4096, limit=C2_NODE_LIMIT, tune=True):
    """The batched leg: `instances` uneven puzzles in one enumeration, its
    launch picked by the autotuner on the workload itself (before timing)."""
    from edge_matching_puzzle_amd.workloads import c2_puzzles
    puzzles = c2_puzzles(instances)
    tuned, trials = {}, None
    if tune:
        from edge_matching_puzzle_amd import autotune
        best, history = autotune.tune(ctx, puzzles, limit, space=BATCHED_SPACE, budget_s=30.0, max_solutions=4,
                                      use_graph=True, log=lambda m: print(m, file=sys.stderr, flush=True))
        tuned = autotune.launch_options(ctx, best)
        trials = [dict(s, nodes_per_s=r) for s, r in history]
        print("batched autotune best %s" % json.dumps(best, sort_keys=True), file=sys.stderr, flush=True)
    for i in range(warmup):
        ctx.enumerate(puzzles, max_solutions=4, node_limit=limit, use_graph=True, **tuned)
        print("batched warmup step %d done" % i, file=sys.stderr, flush=True)
    nodes = launches = 0
    kernel_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        res, st = ctx.enumerate(puzzles, max_solutions=4, node_limit=limit, use_graph=True, **tuned)
        step_nodes = sum(r["n_nodes"] for r in res)
        nodes += step_nodes
        kernel_ms += st["dfs_kernel_ms"]
        launches += st["dfs_launches"]
        print("batched step done: %d nodes, %.1f ms kernel" % (step_nodes, st["dfs_kernel_ms"]),
              file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    done = [r for r in res if r["complete"]]
    times = sorted(r["solve_ms"] for r in done)
    pct = lambda q: times[min(len(times) - 1, int(q * len(times)))] if times else None
    # algorithmic work of the same batch (instrumented run, not timed)
    wres, wst = ctx.enumerate(puzzles, max_solutions=4, node_limit=limit, use_graph=True, count_work=True, **tuned)
    work = (sum(r["n_nodes"] for r in wres), wst["work_bytes"], wst["work_ops"])
    return {
        "value": nodes / elapsed, "unit": "nodes/s", "steps": steps, "warmup": warmup,
        "ms_per_step": elapsed * 1e3 / steps,
        "config": {"workload": "batched enumeration: %d synthetic puzzles, side 5..10, border colours 3..5, "
                               "interior colours 4..12, one enumeration, DFS passes replayed from hipGraphs, "
                               "each instance stopped near %d nodes" % (len(puzzles), limit),
                   "completed": len(done), "cut_at_node_limit": len(res) - len(done),
                   "solve_ms_p50": pct(0.5), "solve_ms_p99": pct(0.99), "nodes_per_step": nodes // steps,
                   "dfs_launches_per_step": launches // steps, "kernel_ms_per_step": kernel_ms / steps,
                   "group_ms_per_step": st["group_ms"], "launch": {"blocks": st["blocks"], "threads": st["threads"]},
                   "tuned": tuned or None, "autotune_trials": trials},
        "roofline": _roofline(nodes, kernel_ms, work, launches),
        "cpu_baseline": cpu,
    }


def bench_c2(args, ctx):
    """The batched workload alone, as a full line."""
    b = bench_batched(ctx, args.steps, args.warmup, getattr(args, "cpu_c2", None), args.instances, args.node_limit,
                      tune=not args.no_autotune)
    return {"metric": METRIC, "value": b["value"], "unit": "nodes/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": b["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u64", "data": "synthetic", "config": b["config"],
            "roofline": b["roofline"], "cpu_baseline": b["cpu_baseline"]}


def bench_c3(args, ctx):
    """C3: glutton_max on 16x16 from hint + 24 spiral moves, push budget."""
    from edge_matching_puzzle_amd.workloads import c3_prefixes, c3_puzzle
    w, h, c, hint = c3_puzzle()
    rows = c3_prefixes(args.prefixes)
    for i in range(args.warmup):
        ctx.glutton_max(w, h, c, prefixes=rows[:256], push_budget=2)
        print("warmup step %d done" % i, file=sys.stderr, flush=True)
    chunk = args.chunk if args.chunk > 0 else len(rows)
    t0 = time.perf_counter()
    pushes = 0
    for _ in range(args.steps):
        # One launch per chunk of prefixes (long runs print as they go).
        res, st, boards = [], {}, []
        for lo in range(0, len(rows), chunk):
            r, b, s = ctx.glutton_max(w, h, c, prefixes=rows[lo:lo + chunk], push_budget=args.push_budget)
            res += r
            boards.append(b)
            for key in ("pushes", "kernel_ms", "evaluations", "score_evaluations", "rounds", "free_slot_sum"):
                st[key] = st.get(key, 0) + s[key]
            st["waves"], st["lds_bytes"] = s["waves"], s["lds_bytes"]
            if chunk < len(rows):
                print("prefixes %d..%d: %d pushes, max depth %d" % (lo, lo + len(r) - 1, s["pushes"],
                                                                    max(x["max_depth"] for x in r)),
                      file=sys.stderr, flush=True)
        boards = np.concatenate(boards)
        pushes += st["pushes"]
        print("step done: %d pushes, %.1f ms kernel, %.3g evaluations" % (st["pushes"], st["kernel_ms"], st["evaluations"]),
              file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    depths = [r["max_depth"] for r in res]
    invalid = _invalid_glutton_boards(w, h, c, rows, res, boards)
    if invalid:
        raise SystemExit("glutton_max returned %d invalid boards (first: prefix %d)" % (len(invalid), invalid[0]))
    return {
        "metric": METRIC, "value": pushes / elapsed, "unit": "glutton nodes (pushes)/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": "C3 glutton_max: 16x16, 5 border / 17 interior colours, seed 3, centre hint + 24 "
                               "spiral moves, %d prefixes, %d pushes each" % (len(rows), args.push_budget),
                   "max_depth": max(depths), "prefixes_depth_ge_200": sum(1 for d in depths if d >= 200),
                   "valid_boards": len(depths),
                   "mean_depth": sum(depths) / len(depths), "evaluations_per_s": st["evaluations"] / (st["kernel_ms"] * 1e-3),
                   "evaluations_per_push": st["evaluations"] / max(st["pushes"], 1),
                   "scoring_share": st["score_evaluations"] / max(st["evaluations"], 1),
                   "removal_rounds_per_push": st["rounds"] / max(st["pushes"], 1),
                   "waves": st["waves"], "lds_bytes_per_wave": st["lds_bytes"]},
        "roofline": _glutton_roofline(w * h, st),
        "cpu_baseline": None if args.no_cpu else _cpu_c3(w, h, c, rows[0]),
    }


def _glutton_roofline(P, st):
    """glutton_kernel against the LDS roofline by the reference's work per
    candidate evaluation (include/CUDA_glutton_max.h:455-714): one mask AND
    per other free position and up to four colour-mask ANDs, each mask
    ceil(4P/64) u64 words -- 8 nw (S_free + 3) bytes, S_free summed over the
    evaluations by the kernel.  The kernel reaches the same choices with far
    fewer reads (slot contexts, affected slots only), so this is the
    reference-equivalent rate."""
    nw = math.ceil(4 * P / 64)
    bytes_ = 8 * nw * (st["free_slot_sum"] + 3 * st["evaluations"])
    achieved = bytes_ / (st["kernel_ms"] * 1e-3) / 1e9 if st["kernel_ms"] else 0.0
    return {"bound": "lds", "achieved": achieved, "peak": LDS_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / LDS_PEAK_GBS, "traffic": None,
            "bytes_per_evaluation": bytes_ / max(st["evaluations"], 1), "kernel_ms": st["kernel_ms"],
            "definition": "reference mask reads per candidate evaluation, 8 nw (S_free + 3) bytes",
            "note": "a fraction above 1 is the reference-equivalent rate: the kernel evaluates a candidate "
                    "from per-slot colour counts instead of one pass over every free position"}


def _invalid_glutton_boards(w, h, c, rows, res, boards):
    """Indices of deepest situations that break the puzzle's edge rules, drop
    a prefix move or disagree with the reported depth / status (checked from
    the puzzle alone, tests/board_check.py)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from board_check import invalid_boards
    import edge_matching_puzzle_amd as emp
    depth = np.array([r["max_depth"] for r in res])
    solved = np.array([r["status"] == emp.GLUTTON_SOLVED for r in res])
    return [int(i) for i in invalid_boards(w, h, c, boards, prefixes=rows, depths=depth, solved=solved)]


def _cpu_c3(w, h, c, row):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bindings as o
    moves = [v for m in row for v in ((int(m) & 0xFF), (int(m) >> 8) & 0xFF, (int(m) >> 16) & 3)]
    t = time.perf_counter()
    r = o.glutton_max(w, h, c, prefix=moves, push_limit=1)
    dt = time.perf_counter() - t
    return {"value": r["pushes"] / dt, "unit": "glutton nodes (pushes)/s", "cores": 1, "kind": "port",
            "sample": "one prefix, one push (CPU oracle, %.1f s)" % dt}


def bench_c4(args, ctx):
    """C4: learned candidate ordering on the 16x16 shape (5 border / 17
    interior colours): train on depth-probe traces of 64 seeds, then compare
    the max depth reached on 16 held-out seeds in equal device time."""
    from edge_matching_puzzle_amd import learned
    t0 = time.perf_counter()
    res = learned.experiment(ctx, time_ms=args.c4_time_ms, log=lambda m: print(m, file=sys.stderr, flush=True))
    elapsed = time.perf_counter() - t0
    res.pop("params")
    per = res.pop("per_seed")
    return {
        "metric": METRIC, "value": res["learned_relabel"]["mean_max_depth"],
        "unit": "mean max depth (pieces placed) per %.0f ms search" % args.c4_time_ms, "n_gpus": 1, "steps": 1,
        "warmup": 0, "ms_per_step": elapsed * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 scorer, u64 search", "data": "synthetic",
        "config": {"workload": "C4 learned ordering: 16x16, 5 border / 17 interior colours, 64 training seeds, "
                               "16 held-out seeds, %.0f ms device time per search" % args.c4_time_ms,
                   "summary": res, "per_seed_max_depth": [{k: (v["max_depth"] if isinstance(v, dict) else v)
                                                           for k, v in r.items() if k != "rank_table_s"} for r in per]},
        "roofline": None, "cpu_baseline": None,
    }


def bench_learned(ctx, time_ms):
    """The learned-ordering leg of the default line, at a reduced size: a
    scorer trained on depth-probe traces of 8 training seeds (64 situations
    each, 5,000-placement probes, 20 epochs), then the max depth reached on 4
    held-out 16x16 seeds in equal device time by the plain order, the trained
    order (ranked kernel, and pieces renumbered by score) and random orders.
    Every order searches the same tree (the ranked and renumbered searches
    are exact reorderings)."""
    from edge_matching_puzzle_amd import learned
    t0 = time.perf_counter()
    res = learned.experiment(ctx, situations_per_seed=64, probe_nodes=5000, epochs=20, time_ms=time_ms,
                             train_seeds=learned.TRAIN_SEEDS[:8], heldout_seeds=learned.HELDOUT_SEEDS[:4],
                             log=lambda m: print(m, file=sys.stderr, flush=True))
    res.pop("params")
    per = res.pop("per_seed")
    orders = ("plain", "learned", "learned_relabel", "random", "random_relabel")
    return {
        "workload": "16x16, 5 border / 17 interior colours; 8 training seeds, 4 held-out seeds, "
                    "%.0f ms device time per search" % time_ms,
        "mean_max_depth": {k: res[k]["mean_max_depth"] for k in orders},
        "nodes_per_s": {k: res[k]["nodes_per_s"] for k in orders},
        "per_seed_max_depth": [{k: r[k]["max_depth"] for k in orders} | {"seed": r["seed"]} for r in per],
        "trace_rows": res["trace_rows"], "heldout_mse": res["heldout_mse"],
        "heldout_mse_of_mean": res["heldout_mse_of_mean"], "loss_first": res["loss_first"],
        "loss_last": res["loss_last"], "seconds": time.perf_counter() - t0,
    }


def hbm_traffic():
    """HBM bytes per dfs_kernel launch of the plain step, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (counters scaled by the reading
    of a plain copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", "dfs_kernel_c1_hbm.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d["bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--workload", choices=["c1", "c2", "c3", "c4"], default="c1")
    ap.add_argument("--c4-time-ms", type=float, default=2000.0, help="C4 device time per search")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after timing, copy 1 GiB with a plain kernel (HBM counter calibration)")
    ap.add_argument("--instances", type=int, default=4096, help="batched instances")
    ap.add_argument("--node-limit", type=int, default=C2_NODE_LIMIT, help="batched nodes per instance")
    ap.add_argument("--batched-steps", type=int, default=2, help="batched leg steps in the default line (0: none)")
    ap.add_argument("--learned-ms", type=float, default=300.0,
                    help="device time per search of the learned leg in the default line (0: none)")
    ap.add_argument("--prefixes", type=int, default=4096, help="C3 prefixes")
    ap.add_argument("--push-budget", type=int, default=16, help="C3 pushes per prefix")
    ap.add_argument("--chunk", type=int, default=0, help="C3 prefixes per launch (0: all in one launch)")
    ap.add_argument("--autotune", action="store_true", help="tune the C1 launch before timing")
    ap.add_argument("--cpu-nodes", type=int, default=600_000_000,
                    help="nodes of the CPU baseline sample (about 10 s on one core)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="batched leg: default launch, no tuning before timing")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import edge_matching_puzzle_amd as emp
    if args.workload in ("c1", "c2") and not args.no_cpu and (args.workload == "c2" or args.batched_steps):
        # the host pool forks, so it runs before this process touches the GPU
        from edge_matching_puzzle_amd.workloads import c2_puzzles
        args.cpu_c2 = _cpu_pool_c2(c2_puzzles(args.instances), args.node_limit)
        print("batched cpu baseline done", file=sys.stderr, flush=True)
    ctx = emp.Context(0)
    if args.workload != "c1":
        line = {"c2": bench_c2, "c3": bench_c3, "c4": bench_c4}[args.workload](args, ctx)
        ctx.close()
        print(json.dumps(line), flush=True)
        return
    w, h, colours = puzzle_c1(args.seed)
    tuned = {}
    if args.autotune:
        from edge_matching_puzzle_amd import autotune
        best, history = autotune.tune(ctx, [(w, h, colours)], 10_000_000_000, budget_s=60,
                                      log=lambda m: print(m, file=sys.stderr, flush=True))
        lanes = ctx.device_info()["compute_units"] * best["waves_per_cu"] * 64
        tuned = dict(waves_per_cu=best["waves_per_cu"], waves_per_block=best["waves_per_block"],
                     steps_per_round=best["steps_per_round"], step_budget=best["step_budget"],
                     target_items=best["target_per_lane"] * lanes)
        print("autotune best %s" % json.dumps(best, sort_keys=True), file=sys.stderr, flush=True)
    for i in range(args.warmup):
        ctx.enumerate([(w, h, colours)], max_solutions=64, **tuned)
        print("warmup step %d done" % i, file=sys.stderr, flush=True)
    nodes = 0
    kernel_ms = 0.0
    launches = 0
    results = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        (r,), st = ctx.enumerate([(w, h, colours)], max_solutions=64, **tuned)
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
    # algorithmic work of the same enumeration (instrumented run, not timed)
    (wr,), wst = ctx.enumerate([(w, h, colours)], max_solutions=64, count_work=True, **tuned)
    if wr["n_nodes"] != r["n_nodes"]:
        raise SystemExit("instrumented run visited %d nodes, timed runs %d" % (wr["n_nodes"], r["n_nodes"]))
    work = (wr["n_nodes"], wst["work_bytes"], wst["work_ops"])

    cpu = None
    if not args.no_cpu:
        rc, dt = cpu_baseline(w, h, colours, args.cpu_nodes)
        cpu = {"value": rc["n_nodes"] / dt, "unit": "nodes/s", "cores": 1, "kind": "port",
               "sample": "first %d nodes of the same DFS (CPU oracle, one thread, %.1f s)" % (rc["n_nodes"], dt)}
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
                   "schedule": "persistent" if st["tickets"] else "passes", "work_items": st["tickets"],
                   "handovers": st["splits"],
                   "lds_tables": bool(st["lds_tables"]), "tuned": tuned or None},
        "cpu_baseline": cpu,
    }
    if args.batched_steps:
        line["batched"] = bench_batched(ctx, args.batched_steps, 1, getattr(args, "cpu_c2", None),
                                        args.instances, args.node_limit, tune=not args.no_autotune)
    if args.learned_ms > 0:
        line["learned"] = bench_learned(ctx, args.learned_ms)
    # The same two peaks measured on this device by stream kernels (after the
    # timed regions): the fractions against them sit beside the published ones.
    pk = ctx.peaks()
    line["roofline"] = _roofline(nodes, kernel_ms, work, launches, pk)
    line["roofline"]["traffic"] = hbm_traffic()
    line["roofline"]["words_per_mask"] = math.ceil(4 * w * h / 64)
    if "batched" in line:
        b = line["batched"]["roofline"]
        b.update({"peak_measured": pk["lds_gbs"], "frac_of_measured": b["achieved"] / pk["lds_gbs"]})
    if args.calibrate_copy:
        print("copy kernel: 1 GiB in %.3f ms" % ctx.stream_copy(1 << 30), file=sys.stderr, flush=True)
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
