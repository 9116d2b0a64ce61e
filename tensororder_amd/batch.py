"""Many independent instances in one schedule (a batch of unsliced plans of uneven size).

Each instance keeps its own compiled join list and its own arena region (placed for level
order).  Step k of the batch runs every join of tree level k of every instance (a join's
level is one more than its children's; joins of one level are independent), as grouped
launches (one per kernel variant) whose blocks locate their join through a block map
(or by binary search over ``block0``), and whose smallest streamed joins run one wave
each, four to a block -- so a batch of N instances needs max_i(height_i) steps rather than
sum_i(#joins_i) launches, and independent small joins share the machine.  Right after its
last level an instance's root is stored into its own accumulator together with the device
clock, which gives per-instance completion times within the batch.
"""

import ctypes

import numpy as np

from . import _lib, schedule
from ._lib import Launch

ALIGN = schedule.ALIGN
KIND_ORDER = schedule.KIND_ORDER
# a grouped launch over at least this many descriptors also carries a block -> descriptor
# map (int32 per block, in aux1), so each block finds its descriptor in one load
MAP_MIN_DESCS = 8


class BatchProgram:
    def __init__(self):
        self.launches = []
        self.info = []
        self.count = 0
        self.stats = {}


def analyse_batch(instances, entry_type, regroup=True):
    """Per instance of a list of (tree, tables) of unsliced plans: its analysed node list
    (joins carry their kernel forms), memory plan and arena peak in bytes."""
    esize = 8 if entry_type == _lib.TO_F64 else 4
    per = []
    for tree, tables in instances:
        nodes = schedule.analyse(tree, tables, regroup)
        if any(n.batched for n in nodes):
            raise ValueError("batched instances must be unsliced plans")
        for n in nodes:
            if not n.is_leaf:
                g = schedule.join_geometry(nodes[n.left].rank, nodes[n.right].rank, n.lmap, n.rmap)
                # kernel forms chosen for fewer launches: the batch is bound by launch count
                g["var"] = schedule.choose_variant(g, 1, esize, launch_bound=True)
                n.geom = g
        peak, scratch = schedule.plan_memory(nodes, tables, 1, esize, "level")
        per.append(dict(nodes=nodes, tables=tables, scratch=scratch, peak=peak,
                        region=(max(peak, ALIGN) + ALIGN - 1) // ALIGN * ALIGN,
                        fixed=sum(t.size for t in tables.values()) * esize + 1024 * len(nodes)))
    return per


def admission_waves(per, budget):
    """Instances (in order) grouped into waves whose arena regions plus fixed bytes (leaf
    tables, descriptors) fit ``budget`` bytes; the waves run one after another in one arena
    of the largest wave's size.  An instance that does not fit alone yields None (the
    caller slices it)."""
    waves, cur, used = [], [], 0
    for i, inst in enumerate(per):
        need = inst["region"] + inst["fixed"]
        if need > budget:
            return None
        if cur and used + need > budget:
            waves.append(cur)
            cur, used = [], 0
        cur.append(i)
        used += need
    if cur:
        waves.append(cur)
    return waves


def wave_arena_bytes(per, wave):
    return sum(per[i]["region"] for i in wave)


def compile_batch(instances, entry_type, device_alloc, regroup=True, per=None, arena=None):
    """``instances``: list of (tree, tables) of unsliced plans.  Returns a BatchProgram.
    ``per`` (from ``analyse_batch``) may be given instead of analysing ``instances`` again,
    and ``arena`` (an object with ``.data_ptr()`` of at least ``wave_arena_bytes`` bytes)
    instead of allocating one."""
    esize = 8 if entry_type == _lib.TO_F64 else 4
    scaled = entry_type == _lib.TO_F32S
    fdtype = np.float64 if esize == 8 else np.float32
    if per is None:
        per = analyse_batch(instances, entry_type, regroup)
    arena_top = 0
    for inst in per:
        inst["base"] = arena_top
        arena_top += inst["region"]

    prog = BatchProgram()
    prog.count = len(per)
    # leaf tables
    chunks, off, builds = [], 0, []
    for inst in per:
        inst["table_off"], inst["exp0"], factor_of = {}, {}, {}
        for n in inst["nodes"]:
            if n.is_leaf and n.tensor not in inst["table_off"]:
                tab = inst["tables"][n.tensor]
                if tab.spec is not None:  # written on the device by the leaf build launch
                    v0, v1, F, e = tab.spec.stored(scaled)
                    builds.append((tab.spec, off, v0, v1))
                    size = tab.size
                else:
                    rows = tab.rows
                    if scaled:
                        rows, F, e = schedule.leaf_encoding(rows)
                    arr = np.ascontiguousarray(rows[:1], dtype=fdtype).ravel()
                    chunks.append((off, arr))
                    size = arr.size
                if scaled:
                    inst["exp0"][n.tensor] = e
                    factor_of[n.tensor] = F
                inst["table_off"][n.tensor] = off
                off += (size * esize + ALIGN - 1) // ALIGN * ALIGN // esize
        # one factor per leaf node (a tensor at two leaves enters the count twice)
        inst["factors"] = [factor_of[n.tensor] for n in inst["nodes"] if n.is_leaf] if scaled else []
    table_host = np.zeros(max(off, 1), dtype=fdtype)
    for o, arr in chunks:
        table_host[o:o + arr.size] = arr
    # per instance: product of its leaves' encoding factors (mantissa, exponent)
    prog.leaf_factors = [schedule.factor_product(inst.get("factors", [])) for inst in per]
    # scale records: constant leaves first (initialised once), then joins (cleared per run)
    nrec, rec_init = 0, []
    if scaled:
        for inst in per:
            for n in inst["nodes"]:
                if n.is_leaf:
                    n.rec, n.rec_stride = nrec, 0
                    rec_init.append((nrec, inst["exp0"][n.tensor]))
                    nrec += 1
        join_rec0 = nrec
        for inst in per:
            for n in inst["nodes"]:
                if not n.is_leaf:
                    n.rec, n.rec_stride = nrec, 0
                    nrec += 1
    prog.table = device_alloc(table_host.nbytes)
    prog.arena = arena if arena is not None else device_alloc(max(ALIGN, arena_top))
    prog.recs = device_alloc(max(1, nrec) * ctypes.sizeof(_lib.ScaleRec))
    prog.accum = device_alloc(len(per) * ctypes.sizeof(_lib.Accum))
    prog.start = device_alloc(64)
    prog.table_host = table_host
    prog.rec_init = rec_init
    prog.rec_count = nrec
    tbase, abase, rbase = prog.table.data_ptr(), prog.arena.data_ptr(), prog.recs.data_ptr()
    rec_size = ctypes.sizeof(_lib.ScaleRec)
    acc_size = ctypes.sizeof(_lib.Accum)

    def recp(n):
        return rbase + n.rec * rec_size if scaled else 0

    for inst in per:
        for n in inst["nodes"]:
            if n.is_leaf:
                n.ptr = tbase + inst["table_off"][n.tensor] * esize
            else:
                n.ptr = abase + inst["base"] + n.buf[0]
            n.bstride = 0

    # group every instance's joins of tree level k into step k.  In the batch graph a step's
    # streamed and split-K joins (with their operand permutes and finishing) run on a side
    # branch beside its permute -> product chain (joins of one level are independent), and
    # the branches join at every step boundary.
    steps = {}
    roots = {}
    for i, inst in enumerate(per):
        nodes = inst["nodes"]
        joins = [n for n in nodes if not n.is_leaf]
        step_of = {j: k for k, grp in enumerate(schedule.join_levels(nodes, False)) for j in grp}
        scratch = {k: {kk: (inst["base"] + blk[0], blk[1]) for kk, blk in v.items()}
                   for k, v in inst["scratch"].items()}
        for n in joins:
            k = step_of[n.idx]
            side = n.geom["var"]["kind"] in ("stream", "splitk")
            for kind, op, variant, lds, st, nblk in schedule.join_records(
                    n, nodes, 1, esize, scaled, abase, scratch.get(n.idx, {}), recp, launch_bound=True):
                # the streamed joins of one step share one launch whatever their form
                # (each descriptor names its own); the step is bound by launch count
                if op == _lib.TO_OP_JOIN_STREAM and variant != _lib.TO_STREAM_WAVE:
                    schedule.check_stream_form(st.kvar)
                    variant = _lib.TO_STREAM_ANY
                key = (side, KIND_ORDER[kind], op, variant)
                if kind == "permute":
                    work = (2 * esize * (int(st.nbatch) << int(st.rank)), 0)
                elif kind in ("stream", "splitk", "gemm"):
                    work = schedule.join_traffic(n, nodes, 1, esize)
                else:
                    work = (0, 0)
                steps.setdefault(k, {}).setdefault(key, []).append((st, nblk, lds, kind, i, work))
        root = nodes[-1]
        rd = _lib.RootDesc(root=root.ptr, rec=recp(root), acc=prog.accum.data_ptr() + i * acc_size)
        roots.setdefault(max(step_of.values()) + 1 if step_of else 0, []).append(rd)

    blob = bytearray()

    def put_bytes(data):
        nonlocal blob
        start = (len(blob) + ALIGN - 1) // ALIGN * ALIGN
        blob += b"\0" * (start - len(blob))
        blob += data
        return start

    def put(structs):
        return put_bytes(b"".join(bytes(st) for st in structs))

    # leaf tables written on the device once, right after upload
    build_off, build_grid = None, 0
    if builds:
        bdescs = []
        for spec, toff, v0, v1 in builds:
            bd = spec.desc(tbase + toff * esize, v0, v1)
            bd.block0 = build_grid
            build_grid += bd.nblocks
            bdescs.append(bd)
        build_off = put(bdescs)

    pending = []
    pending.append((Launch(op=_lib.TO_OP_BATCH_BEGIN, grid=max(1, min(64, -(-len(per) // 256))),
                           aux2=prog.accum.data_ptr(), aux3=prog.start.data_ptr(), iaux0=len(per)), (None, None),
                    ("begin", -1), (0, 0)))
    if scaled and nrec > join_rec0:
        pending.append((Launch(op=_lib.TO_OP_SCALE_CLEAR, grid=max(1, min(1024, -(-(nrec - join_rec0) // 256))),
                               aux0=rbase + join_rec0 * rec_size, iaux0=nrec - join_rec0), (None, None), ("clear", -1),
                        (0, 0)))
    nsteps = max(list(steps) + list(roots)) + 1 if (steps or roots) else 0
    launches_per_kind = {}
    for k in range(nsteps):
        step_start = len(pending)
        for key in sorted(steps.get(k, {})):
            # blocks start in block order: the joins whose blocks run longest go first, so
            # that they do not start last and hold the step open alone
            group = sorted(steps[k][key], key=lambda g: -schedule.block_cost(g[0]))
            structs, block0, lds, wb, wf = [], 0, 0, 0, 0
            waves = key[2] == _lib.TO_OP_JOIN_STREAM and key[3] == _lib.TO_STREAM_WAVE
            for j, (st, nblk, l, kind, i, work) in enumerate(group):
                if waves:
                    # descriptor j runs on wave j % 4 of block j / 4
                    st.block0, st.nblocks = j // 4, 1
                    block0 = j // 4 + 1
                else:
                    st.block0 = block0
                    block0 += nblk
                lds = max(lds, l)
                structs.append(st)
                wb += work[0]
                wf += work[1]
            off = put(structs)
            moff = None
            if len(structs) >= MAP_MIN_DESCS and not waves:
                # block -> descriptor map: one load per block instead of a binary search
                nblks = [g[1] for g in group]
                moff = put_bytes(np.repeat(np.arange(len(structs), dtype=np.int32), nblks).tobytes())
            kind = group[0][3]
            pending.append((Launch(op=key[2], variant=key[3], count=len(structs), grid=block0, lds=lds,
                                   flags=_lib.TO_LAUNCH_SIDE if key[0] else 0), (off, moff), (kind, k), (wb, wf)))
            launches_per_kind[kind] = launches_per_kind.get(kind, 0) + 1
        if k in roots:
            rs = roots[k]
            off = put(rs)
            pending.append((Launch(op=_lib.TO_OP_ROOT_STORE, count=len(rs), grid=max(1, -(-len(rs) // 256))),
                            (off, None), ("root", k), (0, 0)))
        if len(pending) > step_start:
            # the step starts from everything of the previous one, on both branches
            pending[step_start][0].flags |= _lib.TO_LAUNCH_JOIN
    prog.blob_host = np.frombuffer(bytes(blob) or b"\0", dtype=np.uint8).copy()
    prog.blob = device_alloc(prog.blob_host.nbytes)
    bbase = prog.blob.data_ptr()
    prog.build = [] if build_off is None else [
        Launch(op=_lib.TO_OP_LEAF_BUILD, count=len(builds), grid=build_grid, descs=bbase + build_off)]
    # per launch: (kind, step) and the algorithmic (bytes, flops) of its joins
    prog.work = []
    for L, (off, moff), info, work in pending:
        if off is not None:
            L.descs = bbase + off
        if moff is not None:
            L.aux1 = bbase + moff
        prog.launches.append(L)
        prog.info.append(info)
        prog.work.append(work)
    prog.per = per
    prog.stats = dict(instances=len(per), steps=nsteps, launches=len(prog.launches),
                      joins=sum(sum(1 for n in p["nodes"] if not n.is_leaf) for p in per),
                      arena_bytes=arena_top, launches_per_kind=launches_per_kind)
    return prog


def rec_init_bytes(prog):
    n = max(1, prog.rec_count)
    recs = (_lib.ScaleRec * n)()
    one = int(np.array([1.0], dtype=np.float32).view(np.uint32)[0])
    for i, e in prog.rec_init:
        recs[i].exp = e
        recs[i].maxbits = one
    return bytes(recs)
