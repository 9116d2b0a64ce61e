"""Benchmark of the contraction executor on one MI355X.

A step is one full contraction of the workload instance (all of its slices), with the
plan compiled and its inputs resident in HBM before the timed region; the step ends with
the count read back to the host.  ``value`` is instances/s of the FP32-scaled executor
over exactly K timed steps after W warm-up steps -- on the plan the learned cost model
picks among the workload's candidate plans where it has them (the default workload, the
4096-slice random 3-CNF), else on the workload's own plan.  Beside it, on the reference
rule's plan: the plain FP64 executor (``speedup_vs_fp64`` = value / its instances/s) and
the FP32-scaled executor, each timed for up to K steps within ``--side-seconds``; the FP64
count is checked against the oracle's stored count or slice prefix, the FP32 counts
against the FP64 count.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Prints one JSON line.  ``roofline`` covers the kernel symbol with the most device time in
one headline step, timed live with HIP events on the executor's stream;
``cpu_baseline`` is the oracle's numpy restatement of the reference executor (same
tensordot calls, same plan) timed on this host at the default BLAS thread count and at
one thread.
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
# float work on the bf16 matrix cores: six bf16 products per float product
X6_PEAK_TFLOPS = BF16_PEAK_TFLOPS / 6


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
            return "k_gemm_x6<%s, %s, %s, %d>" % (S, wide, tr, 256 if v & 32 else 128)
        if v & 8:
            return "k_gemm128_db<%s, %s, %s, %s>" % (T, S, wide, tr)
        if v & 1:
            return "k_gemm128<%s, %s, %s, %s>" % (T, S, wide, tr)
        return "k_gemm<%s, %s, %s>" % (T, S, "true" if v & 2 else "false")
    if op == _lib.TO_OP_JOIN_STREAM:
        if v == _lib.TO_STREAM_WAVE:
            return "k_join_wave<%s, %s>" % (T, S)
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
        if v in (1, 3):  # chunk tiles of 13 / 14 bits (float), 12 / 13 (double)
            return "k_permute_vec_desc<%s, %d>" % (T, (13 if scaled else 12) + (v == 3))
        return "k_permute_desc<%s>" % T
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
            peak = X6_PEAK_TFLOPS
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


def host_description():
    """nproc, the cores this process may run on, and the CPU model of this host."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"nproc": os.cpu_count(), "usable_cores": usable, "cpu_model": model}


def _cpu_leg(doc, budget_s, full):
    """One timing of the oracle executor under the current BLAS thread setting: (instances
    per second, sample description)."""
    import itertools

    from oracle import numpy_oracle
    from tensororder_amd import plan_io
    from tensororder_amd.plan import ALL_SLICERS

    plan = plan_io.plan_from_dict(doc)
    groups0 = len(doc["groups_to_slice"])
    if groups0 == 0 and (plan.memory <= (1 << 23) or full):
        runs = []
        while not runs or (sum(runs) < budget_s and len(runs) < 5):
            t0 = time.perf_counter()
            numpy_oracle.contract_sliced(doc)
            runs.append(time.perf_counter() - t0)
        runs.sort()
        med = runs[len(runs) // 2]
        return 1.0 / med, "%d full contraction(s) of the plan, median %.3f s" % (len(runs), med)
    sliced, how = doc, "the plan's own %d slices" % (1 << groups0)
    if groups0 == 0:
        # an unsliced plan too large for one host pass: slice it like the reference's
        # execution.run would on OOM (greedy_mem) until one slice holds <= 2^23 entries
        while plan.memory > (1 << 23):
            ALL_SLICERS["greedy_mem"].slice_once(plan)
        sliced = plan_io.plan_to_dict(plan)
        how = "the plan sliced by greedy_mem into %d slices (one slice <= 2^23 entries)" % (
            1 << len(sliced["groups_to_slice"]))
    total = 2 ** len(sliced["groups_to_slice"])
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
    return 1.0 / per_instance, ("%s; first %d slices in product order timed (%.1f s, leaf build included), "
                                "scaled to all %d slices = %.1f s per instance"
                                % (how, done, dt, total, per_instance))


def cpu_baseline(doc, budget_s, full=False):
    """The oracle executor (numpy float64 tensordot walk of the same plan, product-order
    FP64 slice sum) on this host, twice: BLAS at its default thread count and at one
    thread (the reference's --thread_limit 1 through threadpoolctl, numpy_apis.py:47-51).

    A bounded sample: one full contraction when it fits the budget, else the first slices
    in product order (the reference's --slice_cutoff), scaled to one instance."""
    import numpy  # noqa: F401  (loads the BLAS library threadpoolctl reports)
    import threadpoolctl

    info = threadpoolctl.threadpool_info()
    threads = max([i.get("num_threads", 1) for i in info] or [1])
    host = host_description()
    value, sample = _cpu_leg(doc, budget_s, full)
    with threadpoolctl.threadpool_limits(limits=1):
        value1, sample1 = _cpu_leg(doc, budget_s, full)
    return {"value": value, "unit": "instances/s", "cores": threads, "kind": "port",
            "sample": "numpy float64 oracle (oracle/numpy_oracle.py), BLAS threads %d: %s" % (threads, sample),
            "single_thread": {"value": value1, "unit": "instances/s", "cores": 1,
                              "sample": "BLAS threads 1: %s" % sample1},
            "blas": [{k: i.get(k) for k in ("internal_api", "num_threads", "version")} for i in info],
            **host}


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


def permute_sweep(rank=26, reps=20):
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
        # the achievable ceiling: a device copy of the same bytes (read once, write once)
        for _ in range(3):
            b.copy_(a)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            b.copy_(a)
        e1.record(stream)
        torch.cuda.synchronize()
        copy_us = e0.elapsed_time(e1) * 1e3 / reps
        copy_gbs = 2 * esize * (1 << rank) / (copy_us * 1e-6) / 1e9
        rows.append({"entry_bytes": esize, "perm": "copy (ceiling)", "us": copy_us, "GB/s": copy_gbs,
                     "frac": copy_gbs / HBM_PEAK_GBS, "correct": True})
        log(rows[-1])
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
                         "frac_of_copy": gbs / copy_gbs, "correct": ok})
            log(rows[-1])
    print(json.dumps({"permute_sweep": rows, "rank": rank}), flush=True)


def gemm_sweep(reps=8):
    """Standalone split products (k_gemm_x6) on the product shapes of the 4096-slice and
    rank-28 workloads, each on three operand fills of the same shape: uniform [0, 1) floats,
    zeros, and floats that bf16 holds exactly (their middle and low bf16 parts are zero).
    The instruction stream is identical across fills; the time differs only through what
    the matrix cores draw (the chip lowers its clock under power load), so the zero fill
    bounds what the kernel's schedule reaches and the uniform fill what random data gets.
    Each time is the fastest of ``reps`` launches after the first."""
    import ctypes

    import torch

    from tensororder_amd import _lib, autotune
    from tensororder_amd.hip_api import HipF32API

    api = HipF32API()
    # (rows log2, columns log2, shared log2, batch, launch variant, workload)
    shapes = [(11, 11, 10, 64, 61, "4096-slice 3-CNF: 256-row float-accumulator product"),
              (11, 10, 12, 64, 31, "4096-slice 3-CNF: wide (double-accumulator) product"),
              (15, 12, 12, 1, 27, "vertex cover n=200: wide product, columns strided"),
              (15, 12, 12, 1, 31, "vertex cover n=200: wide product, columns contiguous")]
    rows = []
    for mlog, nlog, klog, nb, variant, what in shapes:
        na, nbb, nc = nb << (mlog + klog), nb << (nlog + klog), nb << (mlog + nlog)
        a, b, c = api._alloc(4 * na), api._alloc(4 * nbb), api._alloc(4 * nc)
        recs = autotune._recs(api, 3 * nb)
        gd = _lib.GemmDesc()
        gd.a, gd.b, gd.c = a.data_ptr(), b.data_ptr(), c.data_ptr()
        gd.ab, gd.bb, gd.cb = 1 << (mlog + klog), 1 << (nlog + klog), 1 << (mlog + nlog)
        tr = bool(variant & 4)
        gd.ldcm, gd.ldcn = ((1 << nlog), 1) if tr else (1, 1 << mlog)
        rs = ctypes.sizeof(_lib.ScaleRec)
        gd.as_, gd.bs, gd.cs = recs.data_ptr(), recs.data_ptr() + nb * rs, recs.data_ptr() + 2 * nb * rs
        gd.asb = gd.bsb = 1
        gd.mlog, gd.nlog, gd.klog = mlog, nlog, klog
        gd.nbatch, gd.block0, gd.nblocks = nb, 0, 256
        blob = api._alloc(ctypes.sizeof(gd))
        api._upload(blob, bytes(gd))
        L = _lib.Launch(op=_lib.TO_OP_GEMM, variant=variant, count=1, grid=256, descs=blob.data_ptr())
        flops = 2.0 * nb * (1 << (mlog + nlog + klog))
        for fill in ("uniform", "zero", "bf16-exact"):
            with torch.cuda.stream(api._stream):
                for buf, cnt, seed in ((a, na, 1), (b, nbb, 2)):
                    autotune._fill(api, buf, cnt, 4, seed)
                    t = buf[: 4 * cnt].view(torch.float32)
                    if fill == "zero":
                        t.zero_()
                    elif fill == "bf16-exact":
                        t.copy_(t.to(torch.bfloat16).to(torch.float32))
            api._stream.synchronize()
            ms = autotune._time(api, [L], reps=reps)
            tf = flops / (ms * 1e-3) / 1e12
            rows.append({"m": 1 << mlog, "n": 1 << nlog, "k": 1 << klog, "batch": nb,
                         "kernel": kernel_symbol(L, True), "workload": what, "fill": fill, "ms": ms,
                         "TFLOP/s": tf, "frac": tf / X6_PEAK_TFLOPS})
            log(rows[-1])
        del a, b, c, recs, blob
        torch.cuda.empty_cache()
    print(json.dumps({"gemm_sweep": rows, "peak_TFLOP/s": X6_PEAK_TFLOPS}), flush=True)


def timed_steps(exe, steps, warmup, max_s=None):
    """Run ``warmup`` untimed contractions, then time up to ``steps`` contractions (all of
    them unless ``max_s`` is given: then at least one, stopping once ``max_s`` seconds have
    passed), bracketed by device synchronisation.  Returns (steps done, seconds, Result)."""
    import torch

    for _ in range(warmup):
        exe.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        res = exe.run()
        done += 1
        if max_s is not None and time.perf_counter() - t0 >= max_s:
            break
    torch.cuda.synchronize()
    return done, time.perf_counter() - t0, res


def learned_choice(api, cand_path):
    """The plan the trained cost model picks among the workload's candidate plans
    (decompositions x slice specs), and the reference rule's pick (smallest max rank,
    greedy_mem slices: planning.py:159-162)."""
    from tensororder_amd import costmodel, train_costmodel

    model = costmodel.load_default(api)
    t0 = time.perf_counter()
    plans, labels, ref, _doc = train_costmodel.load_candidates(cand_path)
    choice, preds = costmodel.learned_choice(model, api, plans)
    choose_s = time.perf_counter() - t0
    log("learned choice #%d %s predicted %.3f s (reference #%d predicted %.3f s), choice took %.1f s"
        % (choice, labels[choice], preds[choice], ref, preds[ref], choose_s))
    info = {"candidate": choice, "label": labels[choice], "candidates": len(plans), "reference_candidate": ref,
            "predicted_s": preds[choice], "predicted_reference_s": preds[ref], "choice_host_s": choose_s}
    return plans[choice], plans[ref], info


def learned_eval(args):
    """Every candidate plan of the eight held-out formulas (tests/golden/heldout_candidates.json)
    timed in full with the FP32-scaled executor: per formula the learned pick, the
    reference rule's pick, the FLOP rule's pick and the measured best; one JSON line with
    the per-formula results and their medians (speed-up over the reference rule, over the
    FLOP rule, and regret against the measured best)."""
    import numpy as np

    from tensororder_amd import costmodel, train_costmodel
    from tensororder_amd.hip_api import HipF32API

    with open(os.path.join(ROOT, "tests", "golden", "heldout_candidates.json")) as f:
        sets = json.load(f)["sets"]
    api = HipF32API()
    model = costmodel.load_default(api)
    rows = []
    for doc in sets:
        t0 = time.perf_counter()
        r = train_costmodel.evaluate_choices(api, model, doc, measure_all=True)
        m = r["measured_s"]
        r["speedup_vs_reference"] = m["reference"] / m["learned"]
        r["speedup_vs_flops"] = m["flops"] / m["learned"]
        r["regret_vs_best"] = m["learned"] / r["best_s"] - 1.0
        r["reference_regret_vs_best"] = m["reference"] / r["best_s"] - 1.0
        rows.append(r)
        log("%s: %d candidates in %.0f s; learned #%d %.3f s, reference #%d %.3f s, flops #%d %.3f s, best #%d %.3f s"
            % (r["name"], r["candidates"], time.perf_counter() - t0, r["picks"]["learned"], m["learned"],
               r["picks"]["reference"], m["reference"], r["picks"]["flops"], m["flops"], r["best"], r["best_s"]))
    summary = {k: float(np.median([r[k] for r in rows]))
               for k in ("speedup_vs_reference", "speedup_vs_flops", "regret_vs_best", "reference_regret_vs_best")}
    summary["learned_faster_than_reference"] = sum(r["speedup_vs_reference"] > 1 for r in rows)
    summary["learned_is_best"] = sum(r["picks"]["learned"] == r["best"] for r in rows)
    print(json.dumps({"learned_eval": rows, "median": summary, "formulas": len(rows),
                      "executor": "fp32-scaled", "host": host_description()}), flush=True)


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
        if args.mem_limit:
            api.add_argument("mem_limit", args.mem_limit)
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
        launches = exe.launches
        arr = _lib.launch_array(launches)
        ms = (ctypes.c_float * len(launches))()
        _lib.check(api._lib.to_run_timed(ctypes.addressof(arr), len(launches), api._entry_type,
                                         ctypes.c_void_p(api._stream.cuda_stream), ms), "to_run_timed")
        scaled = api._entry_type == _lib.TO_F32S
        symbols = {}
        for L, (b, fl), t in zip(launches, exe.work, ms):
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
                     "launches_per_step": len(launches), "admission_waves": len(exe.progs),
                     "steps_per_batch": sum(p.stats["steps"] for p in exe.progs),
                     "device_ms_per_step": float(sum(ms)),
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
    ap.add_argument("--workload", default="cnf70_factor_s12")
    ap.add_argument("--side-seconds", type=float, default=20.0,
                    help="time budget of each comparison leg (FP64, FP32 on the reference-rule plan)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--mem-limit", type=int, default=None,
                    help="device byte budget of the executor (a batch over it runs in admission waves)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-full", action="store_true", help="time the full unsliced oracle contraction")
    ap.add_argument("--fp32-only", action="store_true",
                    help="time only the FP32-scaled executor on the workload's plan (for profiler runs)")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned plan choice where the workload has candidate plans")
    ap.add_argument("--permute-sweep", action="store_true",
                    help="only time standalone bit permutations of rank-26 tensors")
    ap.add_argument("--learned-eval", action="store_true",
                    help="only time every candidate plan of the held-out formulas (learned vs reference vs FLOP rule)")
    ap.add_argument("--gemm-sweep", action="store_true",
                    help="only time the split products on uniform, zero and bf16-exact operands")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device copy of 512 MiB three times (reference for HBM counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.permute_sweep:
        permute_sweep()
        return
    if args.gemm_sweep:
        gemm_sweep()
        return
    if args.learned_eval:
        learned_eval(args)
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
    exp = doc["meta"]["expected"]
    expected = exp.get("float64")
    log("workload", args.workload, "maxrank", doc["meta"]["maxrank"], "groups", len(doc["groups_to_slice"]))
    cand_path = os.path.join(ROOT, "tests", "golden", args.workload.rsplit("_", 2)[0] + "_candidates.json")
    use_learned = os.path.exists(cand_path) and not args.no_learned

    def leg(cls, plan, steps, warmup, max_s=None, kernels=True, limit=None):
        api = cls()
        exe = api.prepare(plan, num_slice_limit=limit)
        done, dt, res = timed_steps(exe, steps, warmup, max_s)
        log(cls.__name__, "steps", done, "s", round(dt, 4), "count", res.value, "log10", res.log10)
        classes = symbols = None
        if kernels:
            classes, symbols = per_kernel_times(api, exe)
            log(cls.__name__, "kernels", {k: (round(v["ms"], 3), v["launches"]) for k, v in classes.items()})
        # release this executor's device memory before the next one sizes its waves
        exe.close()
        del exe
        torch.cuda.empty_cache()
        return dict(steps=done, dt=dt, res=res, classes=classes, symbols=symbols, value=done / dt)

    # headline: the FP32-scaled executor, on the plan the learned cost model picks where the
    # workload has candidate plans (else on the workload's own plan): exactly K timed steps
    learned = None
    if use_learned:
        plan32, ref_plan, learned = learned_choice(HipF32API(), cand_path)
        torch.cuda.empty_cache()
    else:
        plan32 = ref_plan = plan_io.plan_from_dict(doc)
    head = leg(HipF32API, plan32, args.steps, args.warmup)
    if args.fp32_only:
        print(json.dumps({"workload": args.workload, "executor": "fp32-scaled", "steps": args.steps,
                          "ms_per_step": head["dt"] / args.steps * 1e3, "kernel_symbols": head["symbols"]}), flush=True)
        return
    # beside it, on the reference rule's plan (the workload's own plan): the plain FP64
    # executor, and the FP32-scaled executor when the headline ran another plan; these
    # legs time up to K steps within a bounded time
    side_warm = min(args.warmup, 1)
    r64 = leg(HipAPI, ref_plan, args.steps, side_warm, max_s=args.side_seconds, kernels=False)
    r32ref = None
    if use_learned:
        r32ref = leg(HipF32API, ref_plan, args.steps, side_warm, max_s=args.side_seconds, kernels=False)
    # FP64 executor on the oracle's slice prefix, where the fixture stores one
    prefix = None
    if "prefix" in exp:
        k = exp["prefix"]["slices"]
        rp = leg(HipAPI, plan_io.plan_from_dict(doc), 1, 0, kernels=False, limit=k)
        prefix = {"slices": k, "oracle": exp["prefix"]["float64"], "fp64": rp["res"].value,
                  "rel_err": abs(rp["res"].value - exp["prefix"]["float64"]) / abs(exp["prefix"]["float64"])}
        log("fp64 prefix", prefix)

    c64 = r64["res"].value
    rel_32_64 = abs(head["res"].value - c64) / abs(c64)
    rel_64_ref = abs(c64 - expected) / abs(expected) if expected else None
    rel_32_ref = abs(head["res"].value - expected) / abs(expected) if expected else None
    classes, symbols = head["classes"], head["symbols"]
    dom_name = max(symbols, key=lambda k: symbols[k]["ms"])
    roofline = roofline_of(dom_name, symbols[dom_name], args.workload)
    # the permute kernel's own HBM roofline (it is HBM-bound whatever dominates)
    perm = {k: v for k, v in symbols.items() if k.startswith("k_permute")}
    perm_roofline = None
    if perm:
        pname = max(perm, key=lambda k: perm[k]["ms"])
        perm_roofline = roofline_of(pname, perm[pname], args.workload)

    value = head["value"]
    executor = "fp32-scaled" + (", plan chosen by the learned cost model" if learned else "")
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
                   "maxrank": plan32.maxrank, "slices": 2 ** len([g for g in plan32.groups_to_slice if g]),
                   "executor": executor},
        "learned": learned,
        "fp64": {"plan": "reference rule", "value": r64["value"], "ms_per_step": r64["dt"] / r64["steps"] * 1e3,
                 "steps": r64["steps"], "warmup": side_warm, "rel_err_vs_oracle": rel_64_ref,
                 "prefix_vs_oracle": prefix},
        "fp32s_reference_plan": None if r32ref is None else {
            "value": r32ref["value"], "ms_per_step": r32ref["dt"] / r32ref["steps"] * 1e3, "steps": r32ref["steps"],
            "speedup_vs_fp64": r32ref["value"] / r64["value"],
            "rel_err_vs_fp64": abs(r32ref["res"].value - c64) / abs(c64)},
        "speedup_vs_fp64": value / r64["value"],
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
