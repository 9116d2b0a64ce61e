"""Benchmark of the contraction executor on one MI355X.

A step is one full contraction of the workload instance (all of its slices), with the
plan compiled and its inputs resident in HBM before the timed region; the step ends with
the count read back to the host.  ``value`` is instances/s of the FP32-scaled executor;
the plain FP64 executor is timed beside it, and both counts are checked against the
expected float64 count stored with the plan fixture.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Prints one JSON line.  ``roofline`` covers the kernel class with the most device time in
one step, timed live with HIP events on the executor's stream; ``cpu_baseline`` is the
oracle's numpy restatement of the reference executor (same tensordot calls, same plan)
timed on this host.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "instances/sec + %roofline; FP32-scaled vs FP64 executor at rel-err ≤1e-6"
WORKLOADS = {
    "vc200_weighted_line": "weighted vertex cover, random 3-regular graph n=200 m=300, "
                           "w(x)~U(0.1,0.9), line-Flow plan, no slicing",
    "cnf70_factor_s12": "random 3-CNF 70 vars / 140 clauses, w~U(0.1,0.9), factor-Flow plan sliced on "
                        "12 groups (4096 slices, max rank 25)",
    "vc90_weighted_line": "weighted vertex cover, random 3-regular graph n=90, line-Flow plan",
    "vc_batch512": "mixed batch of 512 weighted vertex-cover instances, random 3-regular graphs with even "
                   "n~U[30,150], line-Flow plans of max rank 6-22, one HIP graph per batch",
}
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0  # dense (no sparsity)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_symbol(L, scaled):
    """Name of the HIP kernel one launch record runs, as a profiler lists it."""
    from tensororder_amd import _lib

    T = "float" if scaled else "double"
    S = "true" if scaled else "false"
    op, v = L.op, L.variant
    if op == _lib.TO_OP_GEMM:
        wide, tr = "true" if v & 2 else "false", "true" if v & 4 else "false"
        if v & 16:
            return "k_gemm_x6<%s, %s, %s>" % (S, wide, tr)
        if v & 8:
            return "k_gemm128_db<%s, %s, %s, %s>" % (T, S, wide, tr)
        if v & 1:
            return "k_gemm128<%s, %s, %s, %s>" % (T, S, wide, tr)
        return "k_gemm<%s, %s, %s>" % (T, S, "true" if v & 2 else "false")
    if op == _lib.TO_OP_JOIN_STREAM:
        if v & _lib.TO_STREAM_ANY:
            return "k_join_stream_any<%s, %s>" % (T, S)
        if v & _lib.TO_STREAM_ROWS:
            return "k_join_rows<%s, %s>" % (T, S)
        if v & _lib.TO_STREAM_VEC:
            return "k_join_stream_vec<%s, %s, %d>" % (T, S, v & ~_lib.TO_STREAM_VEC)
        return "k_join_stream<%s, %s, %d>" % (T, S, v)
    if op == _lib.TO_OP_JOIN_FINISH:
        return "%s<%s, %s>" % ("k_join_finish_wide" if v == 1 else "k_join_finish", T, S)
    if op == _lib.TO_OP_PERMUTE:
        return "%s<%s>" % ("k_permute_vec_desc" if v == 1 else "k_permute_desc", T)
    names = {_lib.TO_OP_LEAF_GATHER: "k_leaf_gather", _lib.TO_OP_JOIN_SPLITK: "k_join_splitk",
             _lib.TO_OP_JOIN_FINISH: "k_join_finish", _lib.TO_OP_ROOT_REDUCE: "k_root_reduce",
             _lib.TO_OP_ROOT_STORE: "k_root_store"}
    if op in names:
        return "%s<%s, %s>" % (names[op], T, S)
    return {_lib.TO_OP_SCALE_CLEAR: "k_scale_clear", _lib.TO_OP_WAVE_ADVANCE: "k_wave_advance",
            _lib.TO_OP_STEP_BEGIN: "k_step_begin", _lib.TO_OP_BATCH_BEGIN: "k_batch_begin"}.get(op, "op%d" % op)


def measured_traffic(workload, name):
    """HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of known size)
    of one kernel on one workload, from the committed rocprofv3 counter summary, or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f).get(workload, {}).get(name)
    if rec is None:
        return None
    return rec["fetch_bytes_per_launch"] + rec["write_bytes_per_launch"]


def roofline_of(name, k, workload=None):
    """Roofline object of one kernel symbol from its summed launches of one step."""
    if k["flops"] and k["flops"] / max(k["bytes"], 1) > 20 and name.startswith("k_gemm"):
        achieved = k["flops"] / (k["ms"] * 1e-3) / 1e12 if k["ms"] > 0 else 0.0
        out = {"bound": "mfma", "achieved": achieved, "unit": "TFLOP/s", "kernel": name}
        if name.startswith("k_gemm_x6"):
            # float products on the bf16 matrix cores: each float multiply-add is six bf16
            # ones, so the float flops are priced against the dense bf16 peak / 6
            peak = BF16_PEAK_TFLOPS / 6
            out["peak_basis"] = "dense bf16 peak %.0f TFLOP/s / 6 bf16 products per float product" % BF16_PEAK_TFLOPS
        else:
            peak = FP32_PEAK_TFLOPS if "<float" in name else FP64_PEAK_TFLOPS
        out.update({"peak": peak, "frac": achieved / peak, "traffic": measured_traffic(workload, name),
                    "launches_per_step": k["launches"], "kernel_ms_per_step": k["ms"],
                    "avg_launch_us": 1e3 * k["ms"] / max(k["launches"], 1), "algorithmic_flops_per_step": k["flops"]})
        return out
    achieved = k["bytes"] / (k["ms"] * 1e-3) / 1e9 if k["ms"] > 0 else 0.0
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(workload, name), "kernel": name,
            "launches_per_step": k["launches"],
            "kernel_ms_per_step": k["ms"], "avg_launch_us": 1e3 * k["ms"] / max(k["launches"], 1),
            "algorithmic_bytes_per_step": k["bytes"]}


def per_kernel_times(api, exe):
    """Device ms of every launch of one step, with algorithmic bytes and flops, summed by
    kernel class and by kernel symbol (a symbol is one template instance)."""
    import ctypes

    from tensororder_amd import _lib

    prog = exe.prog
    scaled = api._entry_type == _lib.TO_F32S
    launches = list(exe.step_launches)
    info = [("begin", None)] + list(prog.prologue_info)
    work = [(0, 0)] + list(prog.prologue_work)
    for _ in range(exe.nwaves):
        launches += exe.wave_launches
        info += prog.wave_info
        work += prog.wave_work
    arr = _lib.launch_array(launches)
    ms = (ctypes.c_float * len(launches))()
    _lib.check(api._lib.to_run_timed(ctypes.addressof(arr), len(launches), api._entry_type,
                                     ctypes.c_void_p(api._stream.cuda_stream), ms), "to_run_timed")
    classes, symbols = {}, {}
    for L, (kind, _nidx), (b, f), t in zip(launches, info, work, ms):
        for key, table in ((kind, classes), (kernel_symbol(L, scaled), symbols)):
            c = table.setdefault(key, {"ms": 0.0, "launches": 0, "bytes": 0, "flops": 0})
            c["ms"] += t
            c["launches"] += 1
            c["bytes"] += b
            c["flops"] += f
    return classes, symbols


def cpu_baseline(doc, budget_s, full=False):
    """The oracle executor (numpy float64 tensordot walk of the same plan) on this host.

    When one full contraction would exceed the budget, the plan is sliced with the
    reference's greedy_mem slicer until one slice holds <= 2^23 entries, the first slices
    are timed in product order (the reference's --slice_cutoff) and the time is scaled to
    all slices of the sliced plan, i.e. to one instance.  The sample is stated."""
    import threadpoolctl

    from oracle import numpy_oracle
    from tensororder_amd import plan_io
    from tensororder_amd.plan import ALL_SLICERS

    info = threadpoolctl.threadpool_info()
    threads = max([i.get("num_threads", 1) for i in info] or [1])
    plan = plan_io.plan_from_dict(doc)
    small = plan.memory <= (1 << 23) or full
    if small:
        runs = []
        while not runs or (sum(runs) < budget_s and len(runs) < 5):
            t0 = time.perf_counter()
            numpy_oracle.contract_sliced(doc)
            runs.append(time.perf_counter() - t0)
        runs.sort()
        med = runs[len(runs) // 2]
        return {"value": 1.0 / med, "unit": "instances/s", "cores": threads, "kind": "port",
                "sample": "%d full contraction(s) of the plan with numpy float64 tensordot "
                          "(oracle/numpy_oracle.py), median %.3f s, BLAS threads %d"
                          % (len(runs), med, threads)}
    while plan.memory > (1 << 23):
        ALL_SLICERS["greedy_mem"].slice_once(plan)
    sliced = plan_io.plan_to_dict(plan)
    total = 2 ** len(sliced["groups_to_slice"])
    import itertools

    t0 = time.perf_counter()
    done = 0
    leaves = [numpy_oracle.build_leaf(d) for d in sliced["tensors"]]
    groups = numpy_oracle._slice_groups(sliced)
    for a in itertools.product((0, 1), repeat=len(groups)):
        numpy_oracle.walk(sliced, numpy_oracle.slice_leaves(sliced, leaves, a, groups))
        done += 1
        if time.perf_counter() - t0 >= budget_s or done == total:
            break
    dt = time.perf_counter() - t0
    per_instance = dt / done * total
    return {"value": 1.0 / per_instance, "unit": "instances/s", "cores": threads, "kind": "port",
            "sample": "plan sliced by greedy_mem into %d slices (one slice <= 2^23 entries); first %d "
                      "slices timed with the numpy float64 oracle (%.1f s), scaled to all %d slices = "
                      "%.1f s per instance; BLAS threads %d" % (total, done, dt, total, per_instance, threads)}


def copy_calibration():
    """A plain device copy of known size: 512 MiB read, 512 MiB written per launch.  The
    profiler's FETCH_SIZE / WRITE_SIZE of this launch give the counters' scale."""
    import torch

    n = 1 << 27
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n}), flush=True)


def permute_sweep(rank=26, reps=10):
    """Standalone bit-permutation throughput (to_permute) on rank-``rank`` tensors: for each
    entry size and permutation, algorithmic bytes (read once + write once) per second."""
    import ctypes
    import random

    import torch

    from tensororder_amd import _lib

    lib = _lib.load()
    rng = random.Random(0)
    perms = {
        "reverse": list(range(rank))[::-1],
        "rotate_half": [(i + rank // 2) % rank for i in range(rank)],
        "low8_to_top": list(range(8, rank)) + list(range(8)),
        "random": rng.sample(range(rank), rank),
    }
    rows = []
    stream = torch.cuda.current_stream()
    for esize, dt in ((4, torch.float32), (8, torch.float64)):
        a = torch.rand(1 << rank, dtype=dt, device="cuda")
        b = torch.empty_like(a)
        for name, perm in perms.items():
            p = (ctypes.c_int32 * rank)(*perm)

            def go():
                _lib.check(lib.to_permute(ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()), rank,
                                          ctypes.addressof(p), esize, 1, ctypes.c_void_p(stream.cuda_stream)),
                           "to_permute")

            go()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                go()
            e1.record(stream)
            torch.cuda.synchronize()
            us = e0.elapsed_time(e1) * 1e3 / reps
            gbs = 2 * esize * (1 << rank) / (us * 1e-6) / 1e9
            # numpy.transpose semantics: output axis k is input axis perm[k]
            j = torch.arange(1 << rank, device="cuda")
            src = torch.zeros_like(j)
            for k in range(rank):
                src |= ((j >> (rank - 1 - k)) & 1) << (rank - 1 - perm[k])
            ok = bool(torch.equal(b, a[src]))
            del j, src
            rows.append({"entry_bytes": esize, "perm": name, "us": us, "GB/s": gbs, "frac": gbs / HBM_PEAK_GBS,
                         "correct": ok})
            log(rows[-1])
    print(json.dumps({"permute_sweep": rows, "rank": rank}), flush=True)


def learned_choice_run(api, cand_path, args, fp64_value):
    """The FP32-scaled executor on the plan the trained cost model picks among the
    workload's candidate plans (decompositions x slice specs); the reference rule's pick is
    the workload's own plan, timed above."""
    import torch

    from tensororder_amd import costmodel, train_costmodel

    model = costmodel.load_default(api)
    t0 = time.perf_counter()
    plans, labels, ref, doc = train_costmodel.load_candidates(cand_path)
    choice, preds = costmodel.learned_choice(model, api, plans)
    choose_s = time.perf_counter() - t0
    log("learned choice #%d %s predicted %.3f s (reference #%d predicted %.3f s), choice took %.1f s"
        % (choice, labels[choice], preds[choice], ref, preds[ref], choose_s))
    exe = api.prepare(plans[choice])
    for _ in range(args.warmup):
        exe.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = exe.run()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    exe.close()
    log("learned steps", args.steps, "s", round(dt, 4), "count", res.value)
    return {"value": args.steps / dt, "ms_per_step": dt / args.steps * 1e3, "candidate": choice,
            "label": labels[choice], "candidates": len(plans), "reference_candidate": ref,
            "predicted_s": preds[choice], "predicted_reference_s": preds[ref], "choice_host_s": choose_s,
            "rel_err_vs_fp64_reference_plan": abs(res.value - fp64_value) / abs(fp64_value)}


def batch_bench(args):
    """The mixed batch: every instance of the batch fixture contracted by one grouped
    schedule (one HIP graph launch per step); value = instances/s of the FP32-scaled
    executor, with per-instance completion times (device clock, from the batch start)."""
    import ctypes

    import numpy as np
    import torch

    from tensororder_amd import _lib, plan_io
    from tensororder_amd.hip_api import HipAPI, HipF32API

    with open(os.path.join(ROOT, "tests", "golden", args.workload + ".json")) as f:
        doc = json.load(f)
    t0 = time.perf_counter()
    plans, expected = plan_io.batch_plans(doc)
    log("batch", len(plans), "plans rebuilt in %.1f s" % (time.perf_counter() - t0))
    out = {}
    for name, cls in (("fp32s", HipF32API), ("fp64", HipAPI)):
        api = cls()
        exe = api.prepare_batch(plans)
        for _ in range(args.warmup):
            exe.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            res, lat = exe.run()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        errs = [abs(r.value - e) / abs(e) for r, e in zip(res, expected)]
        rel = max(errs)
        worst = sorted(range(len(errs)), key=lambda i: -errs[i])[:5]
        lat_ms = np.asarray(lat) * 1e3
        prog = exe.prog
        arr = _lib.launch_array(prog.launches)
        ms = (ctypes.c_float * len(prog.launches))()
        _lib.check(api._lib.to_run_timed(ctypes.addressof(arr), len(prog.launches), api._entry_type,
                                         ctypes.c_void_p(api._stream.cuda_stream), ms), "to_run_timed")
        scaled = api._entry_type == _lib.TO_F32S
        symbols = {}
        for L, (b, fl), t in zip(prog.launches, prog.work, ms):
            c = symbols.setdefault(kernel_symbol(L, scaled), {"ms": 0.0, "launches": 0, "bytes": 0, "flops": 0})
            c["ms"] += t
            c["launches"] += 1
            c["bytes"] += b
            c["flops"] += fl
        out[name] = {"value": len(plans) * args.steps / dt, "ms_per_step": dt / args.steps * 1e3,
                     "latency_ms_p50": float(np.percentile(lat_ms, 50)),
                     "latency_ms_p99": float(np.percentile(lat_ms, 99)),
                     "rel_err_vs_oracle_max": rel,
                     "rel_err_vs_oracle_median": float(np.median(errs)),
                     "worst_instances": [{"index": i, "n": doc["instances"][i]["n"],
                                          "maxrank": doc["instances"][i]["maxrank"], "rel_err": errs[i]}
                                         for i in worst],
                     "launches_per_step": len(prog.launches),
                     "steps_per_batch": prog.stats["steps"], "device_ms_per_step": float(sum(ms)),
                     "kernel_symbols": symbols}
        log(name, "batch step %.3f ms, %.1f instances/s, p50 %.3f ms p99 %.3f ms, max rel err %.2e"
            % (out[name]["ms_per_step"], out[name]["value"], out[name]["latency_ms_p50"],
               out[name]["latency_ms_p99"], rel))
        exe.close()
        del exe
        torch.cuda.empty_cache()
    r32 = out["fp32s"]
    dom_name = max(r32["kernel_symbols"], key=lambda k: r32["kernel_symbols"][k]["ms"])
    line = {
        "metric": METRIC, "value": r32["value"], "unit": "instances/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r32["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": args.workload, "description": WORKLOADS.get(args.workload, ""),
                   "instances": len(plans), "executor": "fp32-scaled, grouped batch schedule in one HIP graph"},
        "fp32s": {k: v for k, v in r32.items() if k != "kernel_symbols"},
        "fp64": {k: v for k, v in out["fp64"].items() if k != "kernel_symbols"},
        "fp32s_speedup_vs_fp64": r32["value"] / out["fp64"]["value"],
        "roofline": roofline_of(dom_name, r32["kernel_symbols"][dom_name], args.workload),
        "kernel_symbols": r32["kernel_symbols"],
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="vc200_weighted_line")
    ap.add_argument("--cpu-budget", type=float, default=30.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-full", action="store_true", help="time the full unsliced oracle contraction")
    ap.add_argument("--fp32-only", action="store_true",
                    help="time only the FP32-scaled executor on the workload's plan (for profiler runs)")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned plan choice where the workload has candidate plans")
    ap.add_argument("--permute-sweep", action="store_true",
                    help="only time standalone bit permutations of rank-26 tensors")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device copy of 512 MiB three times (reference for HBM counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.permute_sweep:
        permute_sweep()
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the executor has no multi-GPU path in this build "
              "(slices would shard across GPUs with one FP64 partial sum each); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    if args.workload.startswith("vc_batch"):
        batch_bench(args)
        return

    import torch

    from tensororder_amd import plan_io
    from tensororder_amd.hip_api import HipAPI, HipF32API

    path = os.path.join(ROOT, "tests", "golden", args.workload + ".json")
    doc = json.load(open(path))
    expected = doc["meta"]["expected"].get("float64")
    log("workload", args.workload, "maxrank", doc["meta"]["maxrank"], "groups", len(doc["groups_to_slice"]))

    results = {}
    executors = (("fp32s", HipF32API),) if args.fp32_only else (("fp32s", HipF32API), ("fp64", HipAPI))
    for name, cls in executors:
        api = cls()
        plan = plan_io.plan_from_dict(doc)
        exe = api.prepare(plan)
        for _ in range(args.warmup):
            exe.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            res = exe.run()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        log(name, "steps", args.steps, "s", round(dt, 4), "count", res.value, "log10", res.log10)
        classes, symbols = per_kernel_times(api, exe)
        # release this executor's device memory before the next one sizes its waves
        exe.close()
        del exe
        torch.cuda.empty_cache()
        results[name] = dict(api=api, dt=dt, res=res, classes=classes, symbols=symbols)
        log(name, "kernels", {k: (round(v["ms"], 3), v["launches"]) for k, v in classes.items()})

    if args.fp32_only:
        r32 = results["fp32s"]
        print(json.dumps({"workload": args.workload, "executor": "fp32-scaled", "steps": args.steps,
                          "ms_per_step": r32["dt"] / args.steps * 1e3, "kernel_symbols": r32["symbols"]}), flush=True)
        return
    r32, r64 = results["fp32s"], results["fp64"]
    learned = None
    cand_path = os.path.join(ROOT, "tests", "golden", args.workload.rsplit("_", 2)[0] + "_candidates.json")
    if os.path.exists(cand_path) and not args.no_learned:
        learned = learned_choice_run(r32["api"], cand_path, args, r64["res"].value)
    rel_32_64 = abs(r32["res"].value - r64["res"].value) / abs(r64["res"].value)
    rel_64_ref = abs(r64["res"].value - expected) / abs(expected) if expected else None
    rel_32_ref = abs(r32["res"].value - expected) / abs(expected) if expected else None

    classes, symbols = r32["classes"], r32["symbols"]
    dom_name = max(symbols, key=lambda k: symbols[k]["ms"])
    dom = symbols[dom_name]
    roofline = roofline_of(dom_name, dom, args.workload)
    # the permute kernel's own HBM roofline (it is HBM-bound whatever dominates)
    perm = {k: v for k, v in symbols.items() if k.startswith("k_permute")}
    perm_roofline = None
    if perm:
        pname = max(perm, key=lambda k: perm[k]["ms"])
        perm_roofline = roofline_of(pname, perm[pname], args.workload)

    value = args.steps / r32["dt"] if learned is None else learned["value"]
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "instances/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 / value,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "description": WORKLOADS.get(args.workload, ""),
                   "maxrank": doc["meta"]["maxrank"], "slices": 2 ** len(doc["groups_to_slice"]),
                   "executor": "fp32-scaled" + ("" if learned is None else ", plan chosen by the learned cost model")},
        "fp64": {"value": args.steps / r64["dt"], "ms_per_step": r64["dt"] / args.steps * 1e3,
                 "rel_err_vs_oracle": rel_64_ref},
        "fp32s_reference_plan": {"value": args.steps / r32["dt"], "ms_per_step": r32["dt"] / args.steps * 1e3},
        "fp32s_speedup_vs_fp64": r64["dt"] / r32["dt"],
        "learned": learned,
        "speedup_vs_fp64": value / (args.steps / r64["dt"]),
        "rel_err_fp32s_vs_fp64": rel_32_64,
        "rel_err_fp32s_vs_oracle": rel_32_ref,
        "log10_count": r64["res"].log10,
        "roofline": roofline,
        "permute_roofline": perm_roofline,
        "kernel_classes": classes,
        "kernel_symbols": symbols,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(doc, args.cpu_budget, args.cpu_full)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
