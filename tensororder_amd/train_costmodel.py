"""Collect per-join timing records with the executor, train the cost MLP on the GPU, and
compare its plan choice with the reference rule on a held-out candidate set.

    python -m tensororder_amd.train_costmodel --out DIR [--instances N] [--epochs E]
        [--candidates tests/golden/cnf70_candidates.json] [--measure]

Training instances: seeded random 3-CNFs (56-100 variables, clause ratio 1.6-2.2) planned
with the factor-tree planner over the built-in min-degree decomposition (random ties) and
sliced by a random slice spec of ``planners.SLICE_SPECS``; every fourth instance is a
weighted vertex cover planned with the line planner.  For each, one wave of slices (its
size capped so that a wave stays short) is run with per-launch HIP event timing
(``to_run_timed``) in both entry types; every join contributes one record
(16 features -> log device time, log memory footprint).

Held-out evaluation: the candidate fixture (decompositions x slice specs of one formula);
the reference rule takes the decomposition of smallest unsliced max rank with greedy_mem,
the learned rule the smallest predicted time.  With ``--measure`` both choices are
contracted in full with the FP32-scaled executor and timed; their counts must agree.
"""

import argparse
import ctypes
import json
import math
import os
import time

import numpy as np

from . import _lib, costmodel, decomposition, network, planners, schedule
from .formula import parse_dimacs_text
from .instances import random_3cnf, vertex_cover_cnf
from .plan import ALL_SLICERS

WAVE_FLOPS = 2e13  # one timed wave of a training plan stays below this many flops


def log(*a):
    print(*a, flush=True)


def timed_records(api, plan, wave_flops=WAVE_FLOPS):
    """Records of the first wave of one contraction, the wave capped at ``wave_flops``."""
    total, tree, tables, groups = api._plan_inputs(plan, None)
    nodes = schedule.analyse(tree, tables, api._regroup)
    esize = api.get_entry_size()
    per_slice, _ = costmodel.plan_flops(nodes, tables, esize, api._entry_type)
    cap = max(1, int(wave_flops // max(per_slice, 1.0)))
    saved = api._max_batch
    api._max_batch = max(1, min(saved, 1 << int(math.log2(cap))))
    try:
        W = api.choose_batch(nodes, tables, total)
        exe = api._compile(tree, tables, groups, min(total, W), group_levels=False)
    finally:
        api._max_batch = saved
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
        return costmodel.records_from_timing(exe.prog, info, list(ms), exe.prog.nbatch, esize)
    finally:
        exe.close()


def training_plans(count, seed0=0):
    """(index, plan) of the training corpus (deterministic for a seed)."""
    import random

    rng = np.random.default_rng(seed0)
    for i in range(count):
        if i % 4 == 3:
            n = int(rng.integers(30, 101)) * 2
            f = parse_dimacs_text(vertex_cover_cnf(n, seed0 + i, True), "cachet")
            plan = planners.candidate_plans(network.cnf_count(f), "line", [int(rng.integers(1 << 30))])[0]
            cap = int(rng.integers(16, 25))
            random.seed(seed0 + i)
            ALL_SLICERS["greedy_mem"].slice_until(plan, rank=cap)
        else:
            n = int(rng.integers(56, 101))
            ratio = float(rng.uniform(1.6, 2.2))
            f = parse_dimacs_text(random_3cnf(n, int(round(ratio * n)), seed0 + i, True), "cachet")
            plan = planners.candidate_plans(network.cnf_count(f), "factor", [int(rng.integers(1 << 30))])[0]
            slicer, kw, rseed = planners.SLICE_SPECS[int(rng.integers(len(planners.SLICE_SPECS)))]
            groups = int(rng.integers(6, 15))
            random.seed(rseed + 1000 * i)
            ALL_SLICERS[slicer].slice_until(plan, slices=groups, **kw)
            if plan.maxrank > 27:
                ALL_SLICERS["greedy_mem"].slice_until(plan, rank=27)
        yield i, plan


def load_candidates(path):
    """(plans, labels, reference index, document) of a candidate fixture (a path, or one
    candidate document already loaded)."""
    if isinstance(path, dict):
        doc = path
    else:
        with open(path) as f:
            doc = json.load(f)
    net = network.cnf_count(parse_dimacs_text(doc["dimacs"], doc["weights"]))
    tds = [decomposition.parse_all(t["text"])[0] for t in doc["decompositions"]]
    planner = "line" if doc["planner"].startswith("line") else "factor"
    plans, labels = planners.plan_candidates(net, planner, tds, doc["slices"])
    return plans, labels, planners.reference_candidate(labels), doc


def evaluate_choices(api, model, doc, measure_all=False, reps=1):
    """The plan rules on one candidate set: the learned pick (smallest predicted device
    time), the reference rule's (planning.py:159-162 with greedy_mem slices) and the FLOP
    rule's (smallest estimated FLOPs), each timed in full on ``api`` (and, with
    ``measure_all``, every candidate, which names the measured best).  Returns a dict."""
    from . import costmodel

    plans, labels, ref, doc = load_candidates(doc)
    learned, preds = costmodel.learned_choice(model, api, plans)
    flop = costmodel.flop_choice(plans)
    picks = {"learned": learned, "reference": ref, "flops": flop}
    times, counts = {}, {}
    todo = range(len(plans)) if measure_all else sorted(set(picks.values()))
    for i in todo:
        times[i], counts[i] = measure(api, plans[i], reps=reps)
    out = {"name": doc.get("name"), "candidates": len(plans), "slices": doc["slices"],
           "picks": picks, "labels": {k: labels[v] for k, v in picks.items()},
           "predicted_s": {k: preds[v] for k, v in picks.items()},
           "measured_s": {k: times[v] for k, v in picks.items()}, "counts": {k: counts[v] for k, v in picks.items()}}
    if measure_all:
        best = min(times, key=times.get)
        out.update(best=best, best_label=labels[best], best_s=times[best],
                   all_measured_s=[times[i] for i in range(len(plans))], all_predicted_s=list(preds))
    return out


def measure(api, plan, reps=1, warmup=1):
    """(seconds per full contraction, count) with inputs resident."""
    exe = api.prepare(plan)
    try:
        for _ in range(warmup):
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
    ap.add_argument("--instances", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=400)
    ap.add_argument("--candidates", default=None)
    ap.add_argument("--measure", action="store_true")
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
        log("records: instance %d maxrank %d groups %d rows %d (%.0f s)"
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
    summary = dict(rows=int(len(X)), instances=args.instances, epochs=args.epochs, loss_first=hist[0],
                   loss_last=hist[-1], train_log_time_err_median=float(np.median(err)),
                   train_log_time_err_p90=float(np.percentile(err, 90)))

    if args.candidates:
        api = apis["fp32s"]
        plans, labels, ref, doc = load_candidates(args.candidates)
        choice, preds = costmodel.learned_choice(model, api, plans)
        summary["heldout"] = dict(name=doc["name"], candidates=len(plans), reference=ref, learned=choice,
                                  reference_label=labels[ref], learned_label=labels[choice],
                                  predicted_seconds=preds)
        log("held-out %s: %d candidates, reference #%d predicted %.3f s, learned #%d predicted %.3f s"
            % (doc["name"], len(plans), ref, preds[ref], choice, preds[choice]))
        if args.measure:
            t_ref, v_ref = measure(api, plans[ref])
            t_new, v_new = measure(api, plans[choice])
            rel = abs(v_new - v_ref) / abs(v_ref)
            summary["heldout"].update(reference_seconds=t_ref, learned_seconds=t_new, speedup=t_ref / t_new,
                                      count_rel_diff=rel)
            log("measured: reference %.3f s, learned %.3f s (x%.2f), counts rel diff %.2e"
                % (t_ref, t_new, t_ref / t_new, rel))
    with open(os.path.join(args.out, "costmodel_summary.json"), "w") as f:
        json.dump(summary, f, indent=1)


if __name__ == "__main__":
    main()
