"""Collect per-join timing records with the executor, train the cost MLP on the GPU, and
compare its plan choice with the reference rule on held-out instances.

    python -m tensororder_amd.train_costmodel --out DIR [--instances N] [--epochs E]

Training instances: seeded random 3-CNFs (variables and clause ratio drawn per instance)
planned with the factor-tree planner over the built-in min-degree decomposition and sliced
by greedy_mem to a rank cap; vertex-cover instances with the line planner.  For each, one
wave of slices is run with per-launch HIP event timing (``to_run_timed``) in both entry
types; every join contributes one record (features -> log time, log footprint).

Held-out evaluation: for each held-out formula, candidate plans = 4 decompositions x the
slice sets {greedy_mem, greedy_most} x rank caps; the reference rule takes the smallest
unsliced max rank with greedy_mem at the first cap; the learned rule takes the smallest
predicted time.  Both choices are run in full and timed; their counts must agree.
"""

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

from . import _lib, costmodel, network, planners, schedule
from .formula import parse_dimacs_text
from .instances import random_3cnf, vertex_cover_cnf
from .plan import ALL_SLICERS


def log(*a):
    print(*a, flush=True)


def timed_records(api, plan, max_waves=1):
    """Records of one contraction limited to ``max_waves`` waves."""
    total, tree, tables, groups = api._plan_inputs(plan, None)
    nodes = schedule.analyse(tree, tables, api._regroup)
    W = api.choose_batch(nodes, tables, total)
    limit = min(total, W * max_waves)
    exe = api._compile(tree, tables, groups, limit)
    try:
        launches = list(exe.step_launches)
        info = [("begin", None)] + list(exe.prog.prologue_info)
        for _ in range(exe.nwaves):
            launches += exe.wave_launches
            info += exe.prog.wave_info
        arr = _lib.launch_array(launches)
        ms = (ctypes.c_float * len(launches))()
        _lib.check(api._lib.to_run_timed(ctypes.addressof(arr), len(launches), api._entry_type,
                                         ctypes.c_void_p(api._stream.cuda_stream), ms), "to_run_timed")
        # the prologue runs once; wave joins appear once per wave -- keep the first wave
        seen, keep_info, keep_ms = set(), [], []
        for (kind, nidx), t in zip(info, ms):
            key = (kind, nidx)
            if key in seen:
                continue
            seen.add(key)
            keep_info.append((kind, nidx))
            keep_ms.append(t)
        return costmodel.records_from_timing(exe.prog, keep_info, keep_ms, exe.prog.nbatch,
                                             api.get_entry_size())
    finally:
        exe.close()


def training_plans(count, seed0=0, rank_cap=22):
    rng = np.random.default_rng(seed0)
    for i in range(count):
        if i % 4 == 3:
            n = int(rng.integers(20, 60)) * 2
            f = parse_dimacs_text(vertex_cover_cnf(n, seed0 + i, True), "cachet")
            plan = planners.best_plan(network.cnf_count(f), "line", solver=None)
        else:
            n = int(rng.integers(30, 65))
            ratio = float(rng.uniform(1.6, 2.2))
            f = parse_dimacs_text(random_3cnf(n, int(round(ratio * n)), seed0 + i, True), "cachet")
            plan = planners.candidate_plans(network.cnf_count(f), "factor", [int(rng.integers(1 << 30))])[0]
        ALL_SLICERS["greedy_mem"].slice_until(plan, rank=rank_cap)
        yield i, plan


def candidates(formula_text, tie_seeds=(0, 1, 2, 3), caps=(20, 21, 22, 23)):
    """(plans, labels) for 4 decompositions x {greedy_mem, greedy_most} x rank caps."""
    import random

    f = parse_dimacs_text(formula_text, "cachet")
    net = network.cnf_count(f)
    plans, labels = [], []
    for ts in tie_seeds:
        for slicer in ("greedy_mem", "greedy_most"):
            for cap in caps:
                base = planners.candidate_plans(net, "factor", [ts])[0]
                random.seed(ts * 1000 + cap)
                ALL_SLICERS[slicer].slice_until(base, rank=cap)
                plans.append(base)
                labels.append(dict(decomposition=ts, slicer=slicer, rank_cap=cap,
                                   unsliced_rank=None, groups=len(base.groups_to_slice)))
    return plans, labels


def reference_index(formula_text, labels, plans, tie_seeds=(0, 1, 2, 3), cap=20):
    """The reference rule: the decomposition of smallest unsliced max rank, greedy_mem."""
    f = parse_dimacs_text(formula_text, "cachet")
    net = network.cnf_count(f)
    ranks = [planners.candidate_plans(net, "factor", [ts])[0].maxrank for ts in tie_seeds]
    best_ts = tie_seeds[int(np.argmin(ranks))]
    for i, lab in enumerate(labels):
        if lab["decomposition"] == best_ts and lab["slicer"] == "greedy_mem" and lab["rank_cap"] == cap:
            return i
    return 0


def measure(api, plan, reps=2):
    exe = api.prepare(plan)
    try:
        exe.run()
        t0 = time.perf_counter()
        for _ in range(reps):
            res = exe.run()
        return (time.perf_counter() - t0) / reps, res.value
    finally:
        exe.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--instances", type=int, default=40)
    ap.add_argument("--epochs", type=int, default=400)
    ap.add_argument("--heldout", type=int, default=2)
    ap.add_argument("--heldout-vars", type=int, default=56)
    args = ap.parse_args()
    os.makedirs(args.out, exist_ok=True)
    from .hip_api import HipAPI, HipF32API

    apis = {"fp64": HipAPI(), "fp32s": HipF32API()}
    X, Y = [], []
    t0 = time.time()
    for i, plan in training_plans(args.instances):
        for name, api in apis.items():
            x, y = timed_records(api, plan)
            X.append(x)
            Y.append(y)
        log("records: instance %d maxrank %d groups %d total rows %d (%.0f s)"
            % (i, plan.maxrank, len(plan.groups_to_slice), sum(len(a) for a in X), time.time() - t0))
    X = np.concatenate(X)
    Y = np.concatenate(Y)
    np.savez(os.path.join(args.out, "costmodel_records.npz"), X=X, Y=Y)
    model = costmodel.CostModel(apis["fp32s"], seed=0)
    hist = model.fit(X, Y, epochs=args.epochs, batch=2048, lr=3e-3)
    log("trained: rows %d, loss %.4f -> %.4f (normalised MSE)" % (len(X), hist[0], hist[-1]))
    model.save(os.path.join(args.out, "costmodel_params.npz"))
    pred = model.predict(X)
    err = np.abs(pred[:, 0] - Y[:, 0])
    log("train |log time error| median %.3f p90 %.3f" % (np.median(err), np.percentile(err, 90)))

    summary = dict(rows=int(len(X)), loss_first=hist[0], loss_last=hist[-1],
                   train_log_time_err_median=float(np.median(err)), heldout=[])
    api = apis["fp32s"]
    for h in range(args.heldout):
        text = random_3cnf(args.heldout_vars, 2 * args.heldout_vars, 10000 + h, True)
        plans, labels = candidates(text)
        ref = reference_index(text, labels, plans)
        choice, preds = costmodel.learned_choice(model, api, plans)
        t_ref, v_ref = measure(api, plans[ref])
        t_new, v_new = measure(api, plans[choice])
        rel = abs(v_new - v_ref) / abs(v_ref)
        entry = dict(seed=10000 + h, reference=labels[ref], learned=labels[choice],
                     reference_seconds=t_ref, learned_seconds=t_new, speedup=t_ref / t_new,
                     predicted_reference=preds[ref], predicted_learned=preds[choice], count_rel_diff=rel)
        summary["heldout"].append(entry)
        log("held-out %d: reference %.4f s, learned %.4f s (x%.2f), counts rel diff %.2e"
            % (h, t_ref, t_new, t_ref / t_new, rel))
    with open(os.path.join(args.out, "costmodel_summary.json"), "w") as f:
        json.dump(summary, f, indent=1)


if __name__ == "__main__":
    main()
