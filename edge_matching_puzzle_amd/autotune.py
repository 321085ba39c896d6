"""Autotuner of the enumeration launch: sweeps waves per CU, waves per block,
steps per refill round, the prefix split target (items per lane before the
first DFS pass) and the step budget before a work item splits, on a bounded
run of the workload itself (node limit), and keeps the fastest setting.

The LDS stack needs no depth parameter here: each lane's stack keeps every
level in LDS at one byte per level (boards up to 10x10), so no level spills
to HBM.
"""
import itertools
import json
import time

DEFAULT_SPACE = {
    "waves_per_cu": [12, 16, 20],
    "waves_per_block": [2, 4],
    "steps_per_round": [64, 128, 256],
    "target_per_lane": [2, 4, 8],
    # an item over its budget always requeues what remains (only the split-off
    # second half needs free queue room), so small budgets stay bounded
    "step_budget": [1 << 14, 1 << 16, 1 << 18],
}


def _lanes(ctx, waves_per_cu):
    return ctx.device_info()["compute_units"] * waves_per_cu * 64


def launch_options(ctx, setting):
    """The enumeration keywords of a setting."""
    lanes = _lanes(ctx, setting["waves_per_cu"])
    return dict(waves_per_cu=setting["waves_per_cu"], waves_per_block=setting["waves_per_block"],
                steps_per_round=setting["steps_per_round"], step_budget=setting["step_budget"],
                target_items=setting["target_per_lane"] * lanes)


def measure(ctx, puzzles, node_limit, setting, max_solutions=16, verify=None, **enumerate_kw):
    """Nodes/s of one enumeration under `setting` (wall time of the call);
    verify(setting, results), when given, checks the results."""
    t = time.perf_counter()
    res, st = ctx.enumerate(puzzles, max_solutions=max_solutions, node_limit=node_limit,
                            **launch_options(ctx, setting), **enumerate_kw)
    dt = time.perf_counter() - t
    if verify:
        verify(setting, res)
    nodes = sum(r["n_nodes"] for r in res)
    return nodes / dt, nodes, dt


def tune(ctx, puzzles, node_limit, space=None, budget_s=60.0, log=None, verify=None, **enumerate_kw):
    """Coordinate search: start from the middle of every axis, then sweep one
    axis at a time keeping the best value, until the time budget is used.
    enumerate_kw go to every enumeration (max_solutions, use_graph, ...);
    verify(setting, results) is called on each trial's results.
    Returns (best setting, list of (setting, nodes/s))."""
    space = dict(DEFAULT_SPACE, **(space or {}))
    best = {k: v[len(v) // 2] for k, v in space.items()}
    t0 = time.perf_counter()
    rate, _, _ = measure(ctx, puzzles, node_limit, best, verify=verify, **enumerate_kw)
    history = [(dict(best), rate)]
    best_rate = rate
    for axis in space:
        for value in space[axis]:
            if value == best[axis] or time.perf_counter() - t0 > budget_s:
                continue
            trial = dict(best, **{axis: value})
            r, _, _ = measure(ctx, puzzles, node_limit, trial, verify=verify, **enumerate_kw)
            history.append((dict(trial), r))
            if log:
                log("autotune %s -> %.3g nodes/s" % (json.dumps(trial, sort_keys=True), r))
            if r > best_rate:
                best, best_rate = trial, r
    return best, history


def grid(space=None):
    space = dict(DEFAULT_SPACE, **(space or {}))
    keys = list(space)
    for values in itertools.product(*(space[k] for k in keys)):
        yield dict(zip(keys, values))
