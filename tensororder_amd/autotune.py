"""Autotuner for the dense-product tile shape and the permute grid, cached per shape.

The schedule compiler picks kernel variants by rule (``schedule.choose_variant``).  For
the two kernels whose best launch shape depends on the problem size -- the MFMA product
(64x64 or 128-row output tiles, number of blocks) and the 16-byte permute (number of
blocks) -- this module times every candidate once per shape on synthetic operands of that
shape (HIP events, ``to_run_timed``) and keeps the fastest in a JSON cache:

* product key: entry size, log2 rows, log2 columns, log2 shared range, log2 batch items;
* permute key: entry size, rank, log2 batch items.

``schedule.TUNED`` holds the cache the compiler consults; a shape missing from it keeps
the rule's choice.  ``HipAPI.add_argument("autotune", True)`` tunes every new shape of a
plan before compiling it.  The choice never changes a result: every candidate computes the
same product or permutation (only tiling and grid differ).
"""

import ctypes
import json
import math
import os

import numpy as np

from . import _lib, schedule

DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "autotune.json")
TUNE_BYTES = 1 << 31  # operand bytes of one synthetic timing run (batch items are cut to fit)


gemm_key = schedule.gemm_key
perm_key = schedule.perm_key


def load(path=None):
    path = path or os.environ.get("TENSORORDER_TUNE_CACHE", DEFAULT_CACHE)
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def save(cache, path=None):
    path = path or os.environ.get("TENSORORDER_TUNE_CACHE", DEFAULT_CACHE)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def gemm_candidates(esize, mlog, nlog, klog, nb, wide):
    """(tile variant, block cap) pairs valid for a product shape."""
    out = []
    tn = 6 if wide else 7
    tiles64 = nb << (mlog - 6 + nlog - 6)
    for cap in (2048, 8192):
        out.append(((2 if wide else 0), int(min(tiles64, cap))))
    if mlog >= 7 and nlog >= tn and klog >= schedule.GEMM128_MIN_SHARED[(esize, wide)]:
        tiles128 = nb << (mlog - 7 + nlog - tn)
        for cap in (1024, 2048, 8192):
            out.append((1 | (2 if wide else 0), int(min(tiles128, cap))))
        if nlog >= 7:
            # 128x128 tiles at one block per CU (LDS double-buffered)
            tiles128 = nb << (mlog - 7 + nlog - 7)
            for cap in (256, 512, 1024, 8192):
                out.append((1 | (2 if wide else 0) | 8, int(min(tiles128, cap))))
            if esize == 4 and schedule.GEMM_X6 and (klog >= 7 or not wide):
                # float products on the bf16 matrix cores (one block per CU; 256-row tiles
                # with float accumulators)
                tall = not wide and mlog >= 8
                tiles_x6 = tiles128 >> (1 if tall else 0)
                for cap in (256, 512, 1024, 8192):
                    out.append((1 | (2 if wide else 0) | 8 | 16 | (32 if tall else 0), int(min(tiles_x6, cap))))
    return sorted(set(out))


def _fill(api, buf, count, esize, seed):
    """Uniform [0, 1) operand values generated on the device."""
    torch = api._torch
    g = torch.Generator(device=api._device).manual_seed(seed)
    dt = torch.float64 if esize == 8 else torch.float32
    with torch.cuda.stream(api._stream):
        buf[: count * esize].view(dt).uniform_(0.0, 1.0, generator=g)


def _recs(api, nb):
    recs = (_lib.ScaleRec * nb)()
    one = int(np.array([1.0], dtype=np.float32).view(np.uint32)[0])
    for i in range(nb):
        recs[i].maxbits = one
    dev = api._alloc(ctypes.sizeof(recs))
    api._upload(dev, bytes(recs))
    return dev


def _time(api, launches, reps=3):
    arr = _lib.launch_array(launches * reps)
    ms = (ctypes.c_float * (len(launches) * reps))()
    _lib.check(api._lib.to_run_timed(ctypes.addressof(arr), len(launches) * reps, api._entry_type,
                                     ctypes.c_void_p(api._stream.cuda_stream), ms), "to_run_timed")
    per = [sum(ms[r * len(launches):(r + 1) * len(launches)]) for r in range(reps)]
    return min(per[1:] if reps > 1 else per)


def tune_gemm(api, mlog, nlog, klog, nb):
    """Fastest (tile variant, blocks) for one product shape, timed on random operands.
    Returns (choice, per-batch-item ms of every candidate)."""
    esize = api.get_entry_size()
    scaled = api._entry_type == _lib.TO_F32S
    wide = esize == 4 and klog >= schedule.WIDE_ACCUM_MIN_SHARED
    per_item = ((1 << (mlog + klog)) + (1 << (nlog + klog)) + (1 << (mlog + nlog))) * esize
    nbt = max(1, min(nb, TUNE_BYTES // per_item))
    a = api._alloc(nbt * (1 << (mlog + klog)) * esize)
    b = api._alloc(nbt * (1 << (nlog + klog)) * esize)
    c = api._alloc(nbt * (1 << (mlog + nlog)) * esize)
    _fill(api, a, nbt * (1 << (mlog + klog)), esize, 1)
    _fill(api, b, nbt * (1 << (nlog + klog)), esize, 2)
    recs = _recs(api, 3 * nbt) if scaled else None
    results = {}
    for variant, nblocks in gemm_candidates(esize, mlog, nlog, klog, nbt, wide):
        gd = _lib.GemmDesc()
        gd.a, gd.b, gd.c = a.data_ptr(), b.data_ptr(), c.data_ptr()
        gd.ab, gd.bb, gd.cb = 1 << (mlog + klog), 1 << (nlog + klog), 1 << (mlog + nlog)
        gd.ldcm, gd.ldcn = 1 << nlog, 1
        if scaled:
            rs = ctypes.sizeof(_lib.ScaleRec)
            gd.as_, gd.bs, gd.cs = recs.data_ptr(), recs.data_ptr() + nbt * rs, recs.data_ptr() + 2 * nbt * rs
            gd.asb = gd.bsb = 1
        gd.mlog, gd.nlog, gd.klog = mlog, nlog, klog
        gd.nbatch, gd.block0, gd.nblocks = nbt, 0, nblocks
        blob = api._alloc(ctypes.sizeof(gd))
        api._upload(blob, bytes(gd))
        L = _lib.Launch(op=_lib.TO_OP_GEMM, variant=variant, count=1, grid=nblocks, descs=blob.data_ptr())
        results["%d/%d" % (variant, nblocks)] = _time(api, [L]) / nbt
    best = min(results, key=results.get)
    v, n = (int(x) for x in best.split("/"))
    v &= 59  # the tile shape; the transposed-output bit is chosen by the output layout
    # the block count scales with the batch: keep it relative to the full tile count
    return {"variant": v, "block_cap": n if nbt == nb else None, "ms_per_item": results[best]}, results


def tune_perm(api, rank, nb):
    """Fastest block cap of the 16-byte permute for one rank (a random permutation)."""
    esize = api.get_entry_size()
    if rank < schedule.PERM_VEC_BITS[esize]:
        return None, {}
    nbt = max(1, min(nb, TUNE_BYTES // (2 * esize << rank)))
    src = api._alloc(nbt * (esize << rank))
    dst = api._alloc(nbt * (esize << rank))
    rng = np.random.default_rng(rank)
    out_pos = [int(x) for x in rng.permutation(rank)]
    results = {}
    for cap in (512, 1024, 2048, 4096):
        pd, grid, lds, pvar = schedule.perm_desc(src.data_ptr(), dst.data_ptr(), rank, out_pos, nbt, 1 << rank,
                                                 esize)
        pd.nblocks = grid = int(min(nbt << pd.nrest, cap))
        blob = api._alloc(ctypes.sizeof(pd))
        api._upload(blob, bytes(pd))
        L = _lib.Launch(op=_lib.TO_OP_PERMUTE, variant=pvar, count=1, grid=grid, lds=lds, descs=blob.data_ptr())
        results[str(cap)] = _time(api, [L]) / nbt
    best = min(results, key=results.get)
    return {"block_cap": int(best), "ms_per_item": results[best]}, results


def tune_plan(api, nodes, W, cache):
    """Tune every product and permute shape of an analysed plan missing from ``cache``
    (joins carry their variants).  Returns the number of shapes tuned."""
    esize = api.get_entry_size()
    tuned = 0
    for n in nodes:
        if n.is_leaf:
            continue
        g, v = n.geom, n.geom["var"]
        if v["kind"] != "gemm":
            continue
        nb = W if n.batched else 1
        key = gemm_key(esize, g["fx"], g["fy"], g["s"], nb)
        if key not in cache:
            cache[key], _ = tune_gemm(api, g["fx"], g["fy"], g["s"], nb)
            tuned += 1
        X, Y = (nodes[n.left], nodes[n.right]) if g["x_is_left"] else (nodes[n.right], nodes[n.left])
        for need, op in ((v["perm_x"], X), (v["perm_y"], Y)):
            nbo = W if (n.batched and op.batched) else 1
            pk = perm_key(esize, op.rank, nbo)
            if need and pk not in cache and op.rank >= schedule.PERM_VEC_BITS[esize]:
                choice, _ = tune_perm(api, op.rank, nbo)
                if choice:
                    cache[pk] = choice
                    tuned += 1
    return tuned


def main():
    """Tune every product and permute shape of the given plan fixtures (both entry types)
    and write the cache:  python -m tensororder_amd.autotune FIXTURE.json ... [--out PATH]"""
    import argparse

    from . import plan_io
    from .hip_api import HipAPI, HipF32API

    ap = argparse.ArgumentParser()
    ap.add_argument("fixtures", nargs="+")
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    cache = load(args.out)
    for cls in (HipF32API, HipAPI):
        api = cls()
        for path in args.fixtures:
            with open(path) as f:
                doc = json.load(f)
            if doc.get("format") == plan_io.BATCH_FORMAT:
                plans = plan_io.batch_plans(doc)[0]
            elif doc.get("format") == "tensororder-candidates/1":
                from .train_costmodel import load_candidates

                plans = load_candidates(path)[0]
            else:
                plans = [plan_io.plan_from_dict(doc)]
            shapes = 0
            for plan in plans:
                total, tree, tables, groups = api._plan_inputs(plan, None)
                nodes = schedule.analyse(tree, tables, api._regroup)
                W = api.choose_batch(nodes, tables, total) if groups else 1
                schedule.peak_bytes(nodes, tables, W, api._entry_type)
                shapes += tune_plan(api, nodes, W, cache)
            print("%s %s: %d new shapes, cache %d" % (cls.__name__, os.path.basename(path), shapes, len(cache)),
                  flush=True)
            save(cache, args.out)
            api._torch.cuda.empty_cache()


if __name__ == "__main__":
    main()
