"""Schedule compiler: a contraction tree plus per-leaf slice tables -> launch records.

The reference walks the tree once per slice, calling tensordot on every join
(``tensor_network.pyx:142-156``).  Here the walk is compiled once:

* every join becomes one index-bit-gather kernel launch (``to_join_desc``), with the
  operand that has more free bits streamed and the other staged in LDS; numpy's output
  order (left free axes, then right free axes; ``contraction_tree.pyx:264-282``) is
  produced directly, so no operand is ever permuted;
* slices run in batches ("waves") of W: a tensor depending on a sliced leaf carries a
  batch dimension, every other tensor is computed once in a prologue and broadcast
  (batch stride 0) -- slice-invariant subtrees are not recomputed per slice;
* leaves follow the batched-slice design of ``jax_apis.py:168-251``: each leaf is
  reordered once to ``[slice row, unsliced axes]`` and a wave gathers row
  ``sum_g lookup[g] * assignment[g]`` for each of its slices on the device;
* buffers are placed by simulating the post-order walk with a first-fit allocator, so
  the arena size is the true peak for a given W, and W is the largest batch that fits.
"""

import ctypes
import math

import numpy as np

from . import _lib
from ._lib import JoinDesc, LeafDesc, Launch

ALIGN = 256
STREAM_MAX_S = 12
# order of a join's launches inside a group: operand permutes, products, split-K finishing
KIND_ORDER = {"permute": 0, "stream": 1, "splitk": 1, "gemm": 1, "finish": 2}
LDS_LIMIT = 64 * 1024
MAX_STREAM_BLOCKS = 4096
# one-wave streamed joins (many-instance batch): at most this many products per lane
WAVE_MAX_LANE_PRODUCTS = 32
# products of the many-instance batch with fewer 64x64 tiles than this split K (in parts of
# at least 2^GEMM_SPLIT_MIN_KLOG) to reach it
GEMM_SPLIT_BLOCKS = 256
GEMM_SPLIT_MIN_KLOG = 10
# split-K joins of the many-instance batch: each split walks >= 2^this steps of 256
SPLITK_BATCH_MIN_LOG_STEPS = 4
# the many-instance batch takes the row form for joins of at least this many output bits
ROWS_MIN_BITS_BATCH = 20


class _Arena:
    """First-fit placement of buffers in one linear address space (bytes)."""

    def __init__(self):
        self.free = [(0, 1 << 62)]
        self.peak = 0

    def alloc(self, nbytes):
        nbytes = max(ALIGN, (nbytes + ALIGN - 1) // ALIGN * ALIGN)
        for i, (start, size) in enumerate(self.free):
            if size >= nbytes:
                if size == nbytes:
                    self.free.pop(i)
                else:
                    self.free[i] = (start + nbytes, size - nbytes)
                self.peak = max(self.peak, start + nbytes)
                return start, nbytes
        raise MemoryError("arena exhausted")

    def release(self, block):
        start, size = block
        self.free.append((start, size))
        self.free.sort()
        merged = []
        for s, z in self.free:
            if merged and merged[-1][0] + merged[-1][1] == s:
                merged[-1] = (merged[-1][0], merged[-1][1] + z)
            else:
                merged.append((s, z))
        self.free = merged


class Node:
    __slots__ = ("idx", "is_leaf", "tensor", "free", "left", "right", "lmap", "rmap",
                 "rank", "batched", "parent", "geom", "buf", "ptr", "bstride", "rec", "rec_stride")

    def __init__(self, idx):
        self.idx = idx
        self.parent = -1
        self.buf = None
        self.rec = None
        self.rec_stride = 0


def flatten_tree(tree):
    """Post-order node list (children before parents) of a reference-style tree."""
    nodes, stack = [], []
    for t in tree.iterate_postorder():
        n = Node(len(nodes))
        n.is_leaf = bool(t.is_leaf)
        n.free = list(t.free_edges)
        n.rank = len(n.free)
        if n.is_leaf:
            n.tensor = t.tensor_index
        else:
            n.right = stack.pop()
            n.left = stack.pop()
            n.lmap = list(t.left_edge_map)
            n.rmap = list(t.right_edge_map)
            nodes[n.left].parent = n.idx
            nodes[n.right].parent = n.idx
        nodes.append(n)
        stack.append(n.idx)
    if len(stack) != 1:
        raise ValueError("tree walk did not end in one root")
    return nodes


def _join_props(lfree, rfree):
    """(left_edge_map, right_edge_map, free edges) of a join (contraction_tree.pyx:248-288)."""
    rpos = {e: i for i, e in enumerate(rfree)}
    lset = set(lfree)
    lmap, rmap, free = [], [], []
    for i, e in enumerate(lfree):
        if e in rpos:
            lmap.append(i)
            rmap.append(rpos[e])
        else:
            free.append(e)
    free += [e for e in rfree if e not in lset]
    return lmap, rmap, free


class _T:
    """A mutable tree node used while regrouping."""

    __slots__ = ("leaf", "tensor", "free", "left", "right", "batched")

    def __init__(self, leaf, tensor=None, free=None, left=None, right=None, batched=False):
        self.leaf, self.tensor, self.free = leaf, tensor, free
        self.left, self.right, self.batched = left, right, batched


def _mk_join(left, right):
    _, _, free = _join_props(left.free, right.free)
    return _T(False, free=free, left=left, right=right, batched=left.batched or right.batched)


def regroup_absorptions(nodes, small=8, max_shared=5, big=14):
    """Reassociate chains in which a large tensor absorbs small operands one at a time.

    ``((B . s1) . s2)`` becomes ``B . (s1 . s2)`` when s1 and s2 are small (rank <= small),
    B is large (rank >= big) and the combined small operand keeps rank <= small with at most
    ``max_shared`` bits shared with B.  The large tensor is then read and written once per
    group instead of once per small operand; the value is the same contraction.  Returns a
    new post-order node list (fresh ``Node`` objects with recomputed edge maps)."""
    built = []
    for n in nodes:
        if n.is_leaf:
            built.append(_T(True, tensor=n.tensor, free=list(n.free), batched=n.batched))
            continue
        t = _mk_join(built[n.left], built[n.right])
        while True:
            l, r = t.left, t.right
            bigc, smallc = (l, r) if len(l.free) >= len(r.free) else (r, l)
            if len(smallc.free) > small or len(bigc.free) < big or bigc.leaf:
                break
            a, b = bigc.left, bigc.right
            inner_big, inner_small = (a, b) if len(a.free) >= len(b.free) else (b, a)
            if len(inner_small.free) > small or len(inner_big.free) < big:
                break
            m = _mk_join(inner_small, smallc)
            shared = len(set(m.free) & set(inner_big.free))
            if len(m.free) > small or shared > max_shared:
                break
            t = _mk_join(inner_big, m)
        built.append(t)
    out, index = [], {}
    stack = [(built[-1], False)]
    while stack:
        t, expanded = stack.pop()
        if t.leaf or expanded:
            n = Node(len(out))
            n.is_leaf = t.leaf
            n.free = list(t.free)
            n.rank = len(t.free)
            n.batched = t.batched
            if t.leaf:
                n.tensor = t.tensor
            else:
                n.left = index[id(t.left)]
                n.right = index[id(t.right)]
                n.lmap, n.rmap, _ = _join_props(out[n.left].free, out[n.right].free)
                out[n.left].parent = n.idx
                out[n.right].parent = n.idx
            index[id(t)] = n.idx
            out.append(n)
        else:
            stack.append((t, True))
            stack.append((t.right, False))
            stack.append((t.left, False))
    return out


def join_geometry(lrank, rrank, lmap, rmap):
    """Bit placement of one join in numpy tensordot order.

    Output index = (A free index << b) | (B free index), with A = left, B = right, the
    free indices C-ordered over the unshared axes.  Axis i of a rank-r tensor is bit
    r-1-i.  Returns the X (streamed) / Y roles and bit positions."""
    s = len(lmap)
    fa = [i for i in range(lrank) if i not in set(lmap)]
    fb = [i for i in range(rrank) if i not in set(rmap)]
    a, b = len(fa), len(fb)
    a_free = [lrank - 1 - fa[a - 1 - t] for t in range(a)]
    b_free = [rrank - 1 - fb[b - 1 - t] for t in range(b)]
    a_sh = [lrank - 1 - lmap[s - 1 - t] for t in range(s)]
    b_sh = [rrank - 1 - rmap[s - 1 - t] for t in range(s)]
    assert a_free == sorted(a_free) and b_free == sorted(b_free)
    x_is_left = (a, lrank) >= (b, rrank)
    if x_is_left:
        return dict(x_is_left=True, fx=a, fy=b, s=s, shf=b, shy=0,
                    xf=a_free, xs=a_sh, yf=b_free, ys=b_sh)
    return dict(x_is_left=False, fx=b, fy=a, s=s, shf=0, shy=b,
                xf=b_free, xs=b_sh, yf=a_free, ys=a_sh)


# Kernel-choice constants (settled by measurement on MI355X; each alternative they rule out
# is recorded beside the kernel it concerns).
# largest GEMM output tile edge to use (64 or 128)
GEMM_TILES = 128
# stage streamed-join outputs through LDS when they form one contiguous run per tile
STAGE_OUTPUTS = True
# the MFMA path for dense joins
USE_GEMM = True
# the 16-byte streaming join form
STREAM_VEC = True
# streamed joins of one level moving fewer bytes than this share one launch of the
# mixed-form kernel
STREAM_MERGE_BYTES = 4 << 20
# the row form of streamed joins with long contiguous output rows
STREAM_ROWS = True
# float GEMMs with at least this many shared bits accumulate partial sums in double
WIDE_ACCUM_MIN_SHARED = 11
# split-K operands with at least this many shared bits are permuted first when their
# lowest shared bits are scattered
SPLITK_PERMUTE_MIN_SHARED = 14
GEMM_MIN_FREE = 6
GEMM_MIN_SHARED = 4
GEMM128_MIN_SHARED = {(8, False): 4, (4, False): 5, (4, True): 5}
MAX_GEMM_BLOCKS = 8192
# by rule, wide float products stay on 128x64 tiles at two blocks per CU (measured faster
# on MI355X than 128x128 at one block per CU); the autotuner chooses per shape either way
GEMM_WIDE_DB = False
# float products with 128x128 (or 256x128) tiles on the bf16 matrix cores, each operand
# split into three bf16 parts
GEMM_X6 = True


def _canonical(free_pos, shared_pos):
    """True when an operand is already row-major [free][shared] (shared bits lowest)."""
    s = len(shared_pos)
    return list(shared_pos) == list(range(s)) and list(free_pos) == list(range(s, s + len(free_pos)))


# Autotuned launch shapes (``autotune.py``): key -> choice; shapes not in it keep the rule.
TUNED = {}


def gemm_key(esize, mlog, nlog, klog, nb):
    return "gemm/%d/%d/%d/%d/%d" % (esize, mlog, nlog, klog, int(math.log2(max(1, nb))))


def perm_key(esize, rank, nb):
    return "perm/%d/%d/%d" % (esize, rank, int(math.log2(max(1, nb))))


def stream_variant(v, g):
    """Launch variant of a streamed join: YC, TO_STREAM_VEC | YC, TO_STREAM_ROWS or
    TO_STREAM_WAVE."""
    if v.get("wave"):
        return _lib.TO_STREAM_WAVE
    if v.get("rows"):
        return _lib.TO_STREAM_ROWS
    return v["yc"] | (_lib.TO_STREAM_VEC if v["vec"] else 0)


# the forms a descriptor of a mixed streamed-join launch (TO_STREAM_ANY) may name
STREAM_FORMS = frozenset([yc for yc in (1, 2, 4, 8, 16)] + [_lib.TO_STREAM_VEC | yc for yc in (1, 2, 4, 8, 16)]
                         + [_lib.TO_STREAM_ROWS])


def check_stream_form(kvar):
    """Reject a streamed-join descriptor whose form a TO_STREAM_ANY launch cannot run
    (the kernel would leave that join's output unwritten)."""
    if int(kvar) not in STREAM_FORMS:
        raise ValueError("streamed join: unknown form %#x" % int(kvar))


def choose_variant(g, nb, esize, launch_bound=False):
    """Kernel variant and launch shape for a join with ``nb`` batch items.  With
    ``launch_bound`` (the many-instance batch, whose steps group the joins of all instances
    by kernel variant, one launch each) a small join keeps the general kernel rather than a
    faster form that would add a launch to its step."""
    fx, fy, s = g["fx"], g["fy"], g["s"]
    K = 1 << s
    if USE_GEMM and fx >= GEMM_MIN_FREE and fy >= GEMM_MIN_FREE and s >= GEMM_MIN_SHARED:
        wide = esize == 4 and s >= WIDE_ACCUM_MIN_SHARED
        # 128-row tiles: 128 columns, except the wide (double-accumulating) form at two
        # blocks per CU, which has registers for 64; at one block per CU (bit 3, LDS
        # double-buffered) the wide form keeps 128 columns
        # K slab of the 128-row kernels: 16 doubles, 32 floats
        big = GEMM_TILES == 128 and fx >= 7 and fy >= (6 if wide else 7) and s >= GEMM128_MIN_SHARED[(esize, wide)]
        # float products on the bf16 matrix cores (three-way split, bit 4): 128x128 tiles
        # at one block per CU (bit 3)
        x6 = big and esize == 4 and fy >= 7 and GEMM_X6 and (s >= 7 or not wide)
        db = big and fy >= 7 and ((wide and GEMM_WIDE_DB) or x6)
        cap = MAX_GEMM_BLOCKS
        tuned = TUNED.get(gemm_key(esize, fx, fy, s, nb))
        if tuned is not None and bool(tuned["variant"] & 2) == wide and (big or not tuned["variant"] & 1) \
                and (GEMM_X6 or not tuned["variant"] & 16) and (s >= 7 or not (wide and tuned["variant"] & 16)):
            big = bool(tuned["variant"] & 1)  # the tile shape; bit 2 follows the layout below
            db = bool(tuned["variant"] & 8)
            x6 = bool(tuned["variant"] & 16)
            cap = tuned.get("block_cap") or MAX_GEMM_BLOCKS
        if launch_bound:
            # one product form (64x64 tiles, float or wide accumulators) for every product
            # of a batch step, so that they share a launch
            big = db = x6 = False
            cap = MAX_GEMM_BLOCKS
        # split products with float accumulators take 256-row tiles (a 64x64 part per wave)
        # when the rows allow: fewer LDS bytes and barriers per product than 128 rows
        tall = x6 and not wide and fx >= 8
        tn = 6 if (wide and not db) else 7
        tiles = nb << ((fx - (8 if tall else 7) + fy - tn) if big else (fx - 6 + fy - 6))
        # float outputs whose columns are contiguous (shy == 0) are written 16 bytes at a
        # time from a transposed product
        tr = big and esize == 4 and g["shy"] == 0
        out = dict(kind="gemm", tile=(1 if big else 0) | (2 if wide else 0) | (4 if tr else 0) | (8 if db else 0)
                   | (16 if x6 else 0) | (32 if tall else 0),
                   nblocks=int(min(tiles, cap)),
                   perm_x=not _canonical(g["xf"], g["xs"]), perm_y=not _canonical(g["yf"], g["ys"]))
        if launch_bound and not big:
            # few 64x64 tiles over a long K (a batch step's products run one tile per block
            # through all of K): split K into parts of >= 2^GEMM_SPLIT_MIN_KLOG until the
            # product has GEMM_SPLIT_BLOCKS blocks; a finishing launch adds the parts
            lsplit = 0
            while (tiles << lsplit) < GEMM_SPLIT_BLOCKS and s - lsplit > GEMM_SPLIT_MIN_KLOG:
                lsplit += 1
            if lsplit:
                nout = nb << (fx + fy)
                out.update(lsplit=lsplit, nsplit=1 << lsplit, nout=nout,
                           nblocks=int(min(tiles << lsplit, cap)),
                           fin_blocks=int(min(-(-nout // 256), MAX_STREAM_BLOCKS)))
        return out
    # the many-instance batch: a join whose outputs need at most WAVE_MAX_LANE_PRODUCTS
    # products per lane of one wave runs as one wave, four joins to a block (the kernel
    # holds the 2^s shared offsets one per lane: s <= 6)
    if launch_bound and s <= 6 and (-(-(nb << (fx + fy)) // 64) << s) <= WAVE_MAX_LANE_PRODUCTS:
        return dict(kind="stream", wave=True, yc=1, nych=1, lds=0, stage=0, vec=False, nblocks=1)
    yc = min(16, 1 << fy)
    stage = 0
    # row form: outputs C[f][y] with 2^fy >= 32 contiguous y and few shared bits, written
    # as whole rows (each thread keeps 16 bytes of y and their Y values in registers)
    if STREAM_ROWS and g["shy"] == 0 and g["shf"] == fy and 5 <= fy <= (10 if esize == 4 else 9) \
            and fx >= 6 and s <= 4 and not (launch_bound and fx + fy + math.log2(max(nb, 1)) < ROWS_MIN_BITS_BATCH):
        lr = min(fx, 6, 11 + (2 if esize == 4 else 1) - fy)  # log2 rows per tile (<= 8 passes)
        return dict(kind="stream", yc=16, nych=1, lds=0, stage=0, vec=False, rows=True,
                    nblocks=int(min(nb << (fx - lr), MAX_STREAM_BLOCKS)))
    # 16-byte form: each thread owns 16 bytes of consecutive f (X's lowest bits are free)
    # and stores its 2^fy outputs per f as 16-byte chunks (all y of one f adjacent)
    lvf = 2 if esize == 4 else 1
    if STREAM_VEC and fx >= 8 + lvf and list(g["xf"][:lvf]) == list(range(lvf)) and g["shy"] == 0 \
            and g["shf"] == fy and fy <= 4 and s <= STREAM_MAX_S and not (s >= 8 and (nb << fx) < 32768):
        lds = ((8 * K + 15) & ~15) + (1 << fy) * K * esize
        if lds <= LDS_LIMIT:
            ntiles = nb << (fx - 8 - lvf)
            # runs of >= 64 bytes per thread leave through 4 KB of LDS per wave
            if (16 // esize) * (1 << fy) * esize >= 64:
                lds += 4 * 4096
            return dict(kind="stream", yc=1 << fy, nych=1, lds=lds, stage=0, vec=True,
                        nblocks=int(min(ntiles, MAX_STREAM_BLOCKS)))
    if fx >= 8:
        # the Y slab is held in LDS as the kernel's double accumulator type
        while yc > 1 and 8 * K + yc * K * 8 > LDS_LIMIT:
            yc //= 2
        lds = 8 * K + yc * K * 8
        # outputs of a tile form one contiguous run when the Y-free bits sit below the f
        # bits and one chunk holds all of them: stage that run through LDS
        if STAGE_OUTPUTS and g["shy"] == 0 and g["shf"] == fy and yc == (1 << fy) and fy >= 1 \
                and lds + (256 << fy) * esize <= LDS_LIMIT:
            stage |= 1
            lds += (256 << fy) * esize
        # the X tile (256 f values x all k) is one contiguous run when every shared bit of
        # X lies below bit 8 + s; stage it through LDS when lanes would otherwise stride
        if STAGE_OUTPUTS and s >= 1 and max(g["xs"]) < 8 + s and lds + (256 << s) * esize <= LDS_LIMIT \
                and (256 << s) * esize <= 32 * 1024:
            stage |= 2
            lds += (256 << s) * esize
    else:
        lds = 8 * K
    nych = (1 << fy) // yc
    threads = (nb << fx) * nych
    stream_ok = s <= STREAM_MAX_S and lds <= LDS_LIMIT
    if stream_ok and not (s >= 8 and threads < 32768):
        if fx >= 8:
            ntiles = nb * (1 << (fx - 8)) * nych
        else:
            ntiles = -(-(nb << fx) // 256) * nych
        return dict(kind="stream", yc=yc, nych=nych, lds=lds, stage=stage, vec=False,
                    nblocks=int(min(ntiles, MAX_STREAM_BLOCKS)))
    nout = nb << (fx + fy)
    nsplit = 1
    if s >= 8:
        nsplit = 1 << int(min(s - 8, max(0, math.floor(math.log2(max(1, 4096 // nout))))))
        if launch_bound:
            # a batch step holds many split-K joins: each split walks at least
            # 2^SPLITK_BATCH_MIN_LOG_STEPS steps of 256 shared indices
            nsplit = min(nsplit, 1 << max(0, s - 8 - SPLITK_BATCH_MIN_LOG_STEPS))
    # finishing: one thread per output, or one block per output when it has many partials
    fin_wide = nsplit >= 32
    fin_blocks = min(nout, MAX_STREAM_BLOCKS) if fin_wide else min(-(-nout // 256), MAX_STREAM_BLOCKS)
    # The 256 threads of a block read shared indices k = 0..255 of each operand; an operand
    # whose 8 lowest shared bits are not its 8 lowest bits in order is read 4 bytes per
    # cache line, so a long one is first permuted into [free][shared] form (shared bits
    # lowest, in k order)
    def scattered(spos):
        return s >= (max(18, SPLITK_PERMUTE_MIN_SHARED) if launch_bound else SPLITK_PERMUTE_MIN_SHARED) \
            and list(spos[:8]) != list(range(8))
    return dict(kind="splitk", nsplit=nsplit, nout=nout, nblocks=int(nout * nsplit),
                fin_blocks=int(fin_blocks), fin_variant=1 if fin_wide else 0,
                perm_x=scattered(g["xs"]), perm_y=scattered(g["ys"]))


class LeafSpec:
    """An OR or variable leaf of the CNF network (``tensor_network_constructions.py:44-159``)
    described compactly, so that its table is written on the device (``TO_OP_LEAF_BUILD``)
    rather than built on the host and uploaded.  ``perm`` lists the leaf's axes in table
    order: sliced axes first, then the others (``tensor.py:95-119``)."""

    __slots__ = ("kind", "rank", "perm", "w_neg", "w_pos", "positive", "output_index")

    def __init__(self, kind, rank, perm, w_neg=0.0, w_pos=0.0, positive=(), output_index=None):
        self.kind, self.rank, self.perm = kind, int(rank), [int(a) for a in perm]
        self.w_neg, self.w_pos = w_neg, w_pos
        self.positive = [bool(p) for p in positive]
        self.output_index = output_index

    def values(self):
        """(val0, val1) as the table holds them (see include/tensororder_hip.h), and the
        array of the table's distinct values."""
        if self.kind == _lib.TO_LEAF_VAR:
            if self.rank == 0:
                v = float(self.w_neg + self.w_pos)
                return v, v, np.array([v])
            return float(self.w_neg), float(self.w_pos), np.array([0.0, self.w_neg, self.w_pos], dtype=np.float64)
        return 0.0, 1.0, np.array([0.0, 1.0])

    def stored(self, scaled):
        """(val0, val1, F, e): the values the device writes and, for the FP32-scaled executor,
        the table's encoding factor and exponent -- those ``leaf_encoding`` gives the whole
        table, which depend only on its set of values."""
        v0, v1, vals = self.values()
        if not scaled:
            return v0, v1, 1.0, 0
        enc, F, e = leaf_encoding(np.concatenate([vals, [v0, v1]]))
        return float(enc[-2]), float(enc[-1]), F, e

    def desc(self, dst, v0, v1):
        """The ``to_leaf_build_desc`` of this leaf's table at device address ``dst``."""
        d = _lib.LeafBuildDesc()
        d.dst, d.val0, d.val1 = dst, v0, v1
        d.kind, d.rank = self.kind, self.rank
        r = self.rank
        for q in range(r):  # table bit q (from the lowest) is axis perm[r-1-q]
            d.pos[q] = r - 1 - self.perm[r - 1 - q]
        d.out_bit = -1
        if self.kind == _lib.TO_LEAF_OR:
            d.negmask = sum(1 << (r - 1 - a) for a, p in enumerate(self.positive) if not p)
            if self.output_index is not None:
                d.out_bit = r - 1 - self.output_index
        d.nblocks = max(1, min(1024, -(-(1 << r) // 256)))
        return d


class LeafTable:
    """Per-tensor slice rows: ``rows`` (float64 [nrows, 2^rank]) -- or, for a leaf written
    on the device, its ``spec`` -- and ``coef`` (one integer per slice group:
    row = sum_g coef[g] * assignment[g])."""

    __slots__ = ("rows", "coef", "rank", "spec", "nrows")

    def __init__(self, rows, coef, rank, spec=None):
        self.coef = [int(c) for c in coef]
        self.rank = rank
        self.spec = spec
        if spec is None:
            self.rows = np.ascontiguousarray(rows, dtype=np.float64).reshape(-1, 1 << rank)
            self.nrows = self.rows.shape[0]
        else:
            self.rows = None
            self.nrows = 1 << (spec.rank - rank)

    @property
    def size(self):
        return self.nrows << self.rank

    @property
    def batched(self):
        return any(self.coef)


class Program:
    """A compiled plan: memory layout, descriptor blob and launch records."""

    def __init__(self):
        self.prologue = []
        self.wave = []
        self.arena_bytes = 0
        self.nbatch = 1
        self.root_const = False
        self.stats = {}


def perm_plan(rank, out_pos, max_tile_bits=12):
    """LDS tile of a bit permutation (input bit p -> output bit out_pos[p]): the union U of
    the m lowest input bits and the input bits landing on the m lowest output bits, m as
    large as keeps |U| <= max_tile_bits; a tensor of at least max_tile_bits bits gets
    exactly that many (the lowest input bits not yet in U are added)."""
    ipos_of_out = [0] * rank
    for p, q in enumerate(out_pos):
        ipos_of_out[q] = p
    U = set()
    m = 0
    while m < rank:
        cand = U | {m, ipos_of_out[m]}
        if len(cand) > max_tile_bits:
            break
        U = cand
        m += 1
    for p in range(rank):
        if len(U) >= max_tile_bits:
            break
        U.add(p)
    u_in = sorted(U)
    v_out, v2u = [], []
    for q in range(rank):
        p = ipos_of_out[q]
        if p in U:
            v_out.append(q)
            v2u.append(u_in.index(p))
    rest = [p for p in range(rank) if p not in U]
    return u_in, v_out, v2u, rest, [out_pos[p] for p in rest]


# tile bits of the vectorised permute kernel: 32 KB tiles of 16-byte chunks (variant 1), or
# 64 KB (variant 3) for tensors with more bits than that
PERM_VEC_BITS = {4: 13, 8: 12}


def perm_desc(src, dst, rank, out_pos, nbatch, stride, esize, big_tiles=True):
    """(to_perm_desc, grid, LDS bytes, variant) of a batched bit permutation; variants 1
    and 3 are the vectorised kernel (tensors of at least its tile bits, 16-byte aligned
    buffers) with 32 KB and 64 KB tiles.  ``big_tiles=False`` keeps to 32 KB tiles, so that
    the permutes of a many-instance step share one launch."""
    vec_bits = PERM_VEC_BITS[esize]
    vec = rank >= vec_bits and src % 16 == 0 and dst % 16 == 0 and (stride * esize) % 16 == 0
    big = vec and big_tiles and rank > vec_bits
    u_in, v_out, v2u, rest_in, rest_out = perm_plan(rank, out_pos, (vec_bits + big) if vec else 12)
    d = _lib.PermDesc()
    d.in_, d.out = src, dst
    d.inb = d.outb = stride
    d.rank, d.nu, d.nrest, d.nbatch = rank, len(u_in), len(rest_in), nbatch
    for arr, vals in ((d.u_in, u_in), (d.v_out, v_out), (d.v2u, v2u), (d.rest_in, rest_in), (d.rest_out, rest_out)):
        for i, x in enumerate(vals):
            arr[i] = x
    ntiles = nbatch << len(rest_in)
    d.block0 = 0
    # the chunk kernel holds 2 or 4 blocks per CU (68 or 36 KB of LDS each): by rule 1024
    # blocks, each walking a contiguous range of tiles
    cap = 1024 if vec else 4096
    tuned = TUNED.get(perm_key(esize, rank, nbatch)) if vec else None
    if tuned is not None:
        cap = tuned["block_cap"]
    d.nblocks = int(min(ntiles, cap))
    lds = ((esize << len(u_in)) + 15) // 16 * 16 + 2 * 256 * 8
    return d, d.nblocks, lds, (3 if big else 1) if vec else 0


def leaf_encoding(rows):
    """Float32 encoding of one leaf table for the scaled executor: (stored values as float64
    that float32 holds exactly, FP64 factor F, power-of-two exponent e) with
    rows ~= F * 2^e * stored.

    A table whose nonzero magnitudes take at most two values -- a variable's weight pair
    (w(not x), w(x)), or 0/1 clause entries -- is stored exactly: the two magnitudes become
    the integers a, b < 2^24 of the best rational approximation b/a of their ratio, and F
    carries the rest (relative error of the ratio ~1e-13).  Each leaf node of the tree enters
    each slice's contraction once, so the whole count scales by the product of F over the
    leaf nodes (a tensor at two leaves counts twice), which the host multiplies back in
    FP64.  Other tables are rounded to float32 (F = 1)."""
    rows = np.asarray(rows, dtype=np.float64)
    mags = np.unique(np.abs(rows[rows != 0]))
    F = 1.0
    stored = rows
    if len(mags) == 1:
        F = float(mags[0])
        stored = np.sign(rows)
    elif len(mags) == 2:
        import fractions

        lo, hi = float(mags[0]), float(mags[1])
        r = hi / lo
        limit = int((1 << 24) - 1)
        fr = fractions.Fraction(r).limit_denominator(max(1, int(limit // r)))
        a, b = fr.denominator, fr.numerator
        if 0 < a <= limit and 0 < b <= limit and abs(lo / a * b - hi) <= 1e-12 * hi:
            F = lo / a
            stored = np.where(np.abs(rows) == lo, np.sign(rows) * a, np.sign(rows) * b)
            stored = np.where(rows == 0, 0.0, stored)
    m = float(np.max(np.abs(stored))) if stored.size else 0.0
    e = math.frexp(m)[1] if m > 0 else 0
    return np.ldexp(stored, -e), F, e


def factor_product(factors):
    """Product of FP64 factors as (mantissa in [0.5, 1), int exponent), without overflow."""
    mant, exp = 1.0, 0
    for f in factors:
        fm, fe = math.frexp(f)
        mant *= fm
        exp += fe
        mant, e2 = math.frexp(mant)
        exp += e2
    return mant, exp


def _pos_array(vals):
    arr = _lib.c_i8x32()
    for i, v in enumerate(vals):
        arr[i] = v
    return arr


def _temporaries(arena, n, nodes, W, esize, partial, keep=False):
    """Scratch buffers a join needs while it runs (split-K partial sums, permuted GEMM and
    split-K operands); allocated after the output and released at once, or returned for the
    caller to release (``keep``) when other joins run beside it."""
    v = n.geom["var"]
    blocks = []
    if v["kind"] == "splitk" or v.get("lsplit"):
        blocks.append(("partial", arena.alloc(v["nout"] * v["nsplit"] * 8)))  # double partial sums
    if v["kind"] in ("gemm", "splitk"):
        g = n.geom
        X, Y = (nodes[n.left], nodes[n.right]) if g["x_is_left"] else (nodes[n.right], nodes[n.left])
        for key, op, need in (("perm_x", X, v["perm_x"]), ("perm_y", Y, v["perm_y"])):
            if need:
                nb = W if (n.batched and op.batched) else 1
                blocks.append((key, arena.alloc(nb * (1 << op.rank) * esize)))
    partial[n.idx] = dict(blocks)
    if keep:
        return [blk for _, blk in blocks]
    for _, blk in blocks:
        arena.release(blk)
    return []


def join_levels(nodes, batched):
    """Joins of one phase (``batched`` False: the prologue, True: the wave) grouped by
    level: a join's level is one more than the highest level among its children of the
    same phase (leaves and the other phase's tensors count 0).  Joins of one level are
    independent, so each level can run as grouped launches."""
    level, out = {}, []
    for n in nodes:
        if n.is_leaf or n.batched != batched:
            continue
        lv = 1
        for c in (n.left, n.right):
            if c in level:
                lv = max(lv, level[c] + 1)
        level[n.idx] = lv
        while len(out) < lv:
            out.append([])
        out[lv - 1].append(n.idx)
    return out


def plan_memory(nodes, tables, W, esize, order="post"):
    """Arena placement for batch W when joins run one at a time in post-order
    (``order="post"``) or level by level, all joins of a level at once (``"level"``).
    Returns (arena peak bytes, scratch placements)."""
    arena = _Arena()
    partial = {}

    def run(group, nb, release_child):
        # outputs and scratch of every join of the group coexist with all of its inputs
        temps = []
        for i in group:
            n = nodes[i]
            n.buf = arena.alloc(nb * (1 << n.rank) * esize)
            temps += _temporaries(arena, n, nodes, nb, esize, partial, keep=True)
        for blk in temps:
            arena.release(blk)
        for i in group:
            n = nodes[i]
            for c in (n.left, n.right):
                if release_child(nodes[c]):
                    arena.release(nodes[c].buf)

    def groups(batched):
        if order == "level":
            return join_levels(nodes, batched)
        return [[n.idx] for n in nodes if not n.is_leaf and n.batched == batched]

    # prologue: slice-invariant joins; their outputs read by the wave stay allocated
    for grp in groups(False):
        run(grp, 1, lambda ch: not ch.is_leaf and not ch.batched)
    # wave: batched leaves are gathered together at wave start
    for n in nodes:
        if n.is_leaf and n.batched:
            n.buf = arena.alloc(W * (1 << n.rank) * esize)
    for grp in groups(True):
        run(grp, W, lambda ch: ch.batched)
    return arena.peak, partial


# level-grouped execution is used when its arena peak stays within this factor of the
# post-order peak
LEVEL_GROUPS = True
LEVEL_PEAK_FACTOR = 1.10


def choose_order(nodes, tables, W, esize):
    """(order, arena peak) for a plan whose joins carry their variants."""
    peak_post, _ = plan_memory(nodes, tables, W, esize, "post")
    if not LEVEL_GROUPS:
        return "post", peak_post
    peak_level, _ = plan_memory(nodes, tables, W, esize, "level")
    if peak_level <= LEVEL_PEAK_FACTOR * peak_post:
        return "level", peak_level
    return "post", peak_post


def analyse(tree, tables, regroup=True):
    """Node list with batched flags (depends on a sliced leaf); with ``regroup`` the
    small-operand absorption chains are reassociated (``regroup_absorptions``)."""
    nodes = flatten_tree(tree)
    for n in nodes:
        if n.is_leaf:
            t = tables[n.tensor]
            if t.rank != n.rank:
                raise ValueError("leaf %d: table rank %d != tree rank %d" % (n.tensor, t.rank, n.rank))
            n.batched = t.batched
        else:
            n.batched = nodes[n.left].batched or nodes[n.right].batched
    if nodes[-1].rank != 0:
        raise ValueError("the tree's root has free edges; only full contractions are supported")
    if regroup:
        nodes = regroup_absorptions(nodes)
    return nodes


def peak_bytes(nodes, tables, W, entry_type):
    esize = 8 if entry_type == _lib.TO_F64 else 4
    for n in nodes:
        if not n.is_leaf:
            nb = W if n.batched else 1
            g = join_geometry(nodes[n.left].rank, nodes[n.right].rank, n.lmap, n.rmap)
            g["var"] = choose_variant(g, nb, esize)
            n.geom = g
    return choose_order(nodes, tables, W, esize)[1]


def join_records(n, nodes, W, esize, scaled, abase, scratch, recp, launch_bound=False):
    """Descriptor records of one join, in execution order:
    [(kind, op, variant, lds, descriptor struct, blocks)].  Descriptors carry block0 = 0;
    a grouped launch rebases them.  ``launch_bound`` (the many-instance batch) keeps every
    operand permute on one tile size."""
    g, v = n.geom, n.geom["var"]
    nb = W if n.batched else 1
    X, Y = (nodes[n.left], nodes[n.right]) if g["x_is_left"] else (nodes[n.right], nodes[n.left])
    out = []
    # operands permuted into [free][shared] form (shared bits lowest, in k order) first
    operands = []
    for key, op, fpos, spos in (("perm_x", X, g["xf"], g["xs"]), ("perm_y", Y, g["yf"], g["ys"])):
        nb_op = W if (n.batched and op.batched) else 1
        stride = (1 << op.rank) if (n.batched and op.batched) else 0
        if v.get(key):
            dst = abase + scratch[key][0]
            out_pos = [0] * op.rank
            for t, p in enumerate(spos):
                out_pos[p] = t
            for i, p in enumerate(fpos):
                out_pos[p] = len(spos) + i
            pd, grid, lds, pvar = perm_desc(op.ptr, dst, op.rank, out_pos, nb_op, 1 << op.rank, esize,
                                               big_tiles=not launch_bound)
            out.append(("permute", _lib.TO_OP_PERMUTE, pvar, lds, pd, grid))
            operands.append((dst, stride, op, True))
        else:
            operands.append((op.ptr, stride, op, False))
    if v["kind"] == "gemm":
        (a, ab, xo, _), (b, bb, yo, _) = operands
        gd = _lib.GemmDesc()
        gd.a, gd.b, gd.c = a, b, n.ptr
        gd.ab, gd.bb, gd.cb = ab, bb, n.bstride
        gd.ldcm, gd.ldcn = 1 << g["shf"], 1 << g["shy"]
        if scaled:
            gd.as_, gd.bs, gd.cs = recp(xo), recp(yo), recp(n)
            gd.asb = xo.rec_stride if n.batched else 0
            gd.bsb = yo.rec_stride if n.batched else 0
        gd.mlog, gd.nlog, gd.klog = g["fx"], g["fy"], g["s"]
        gd.nbatch = nb
        gd.block0, gd.nblocks = 0, v["nblocks"]
        if v.get("lsplit"):
            # split K: the parts' partial sums go to scratch, a finishing launch writes C
            assert not v["tile"] & 1
            gd.c, gd.lsplit = abase + scratch["partial"][0], v["lsplit"]
        out.append(("gemm", _lib.TO_OP_GEMM, v["tile"], 0, gd, v["nblocks"]))
        if not v.get("lsplit"):
            return out
    d = JoinDesc()
    d.x, d.y, d.c = X.ptr, Y.ptr, n.ptr
    d.xb = X.bstride if n.batched else 0
    d.yb = Y.bstride if n.batched else 0
    d.cb = n.bstride
    xpos = (g["xf"], g["xs"])
    ypos = (g["yf"], g["ys"])
    if operands[0][3]:  # X permuted: [free][shared], dense per batch item
        d.x, d.xb = operands[0][0], operands[0][1]
        xpos = (list(range(g["s"], g["s"] + g["fx"])), list(range(g["s"])))
    if operands[1][3]:
        d.y, d.yb = operands[1][0], operands[1][1]
        ypos = (list(range(g["s"], g["s"] + g["fy"])), list(range(g["s"])))
    if scaled:
        d.xs, d.ys, d.cs = recp(X), recp(Y), recp(n)
        d.xsb = X.rec_stride if n.batched else 0
        d.ysb = Y.rec_stride if n.batched else 0
    d.fx, d.fy, d.s = g["fx"], g["fy"], g["s"]
    d.shf, d.shy = g["shf"], g["shy"]
    d.nbatch = nb
    d.block0 = 0
    d.xf_pos, d.xs_pos = _pos_array(xpos[0]), _pos_array(xpos[1])
    d.yf_pos, d.ys_pos = _pos_array(ypos[0]), _pos_array(ypos[1])
    if v["kind"] == "gemm":  # the finishing launch of a split-K product
        d.partial = abase + scratch["partial"][0]
        d.nsplit = v["nsplit"]
        d.nblocks = v["fin_blocks"]
        out.append(("finish", _lib.TO_OP_JOIN_FINISH, 0, 0, d, v["fin_blocks"]))
        return out
    if v["kind"] == "stream":
        d.nblocks, d.nych, d.nsplit = v["nblocks"], v["nych"], 1
        d.lnych = v["nych"].bit_length() - 1
        d.stage = int(v["stage"])
        d.kvar = stream_variant(v, g)
        out.append(("stream", _lib.TO_OP_JOIN_STREAM, d.kvar, v["lds"], d, v["nblocks"]))
        return out
    d.partial = abase + scratch["partial"][0]
    d.nsplit = v["nsplit"]
    d.nblocks = v["nblocks"]
    out.append(("splitk", _lib.TO_OP_JOIN_SPLITK, 0, 0, d, v["nblocks"]))
    d2 = JoinDesc.from_buffer_copy(bytes(d))
    d2.nblocks = v["fin_blocks"]
    out.append(("finish", _lib.TO_OP_JOIN_FINISH, v["fin_variant"], 0, d2, v["fin_blocks"]))
    return out


def block_cost(st):
    """Rough work of one block of a descriptor, to start the longest blocks of a grouped
    launch first: its shared-index loop length times the tiles it walks."""
    if isinstance(st, _lib.GemmDesc):
        tiles = ((st.nbatch << max(0, st.mlog + st.nlog - 12)) << st.lsplit) // max(1, st.nblocks)
        return (1 << (st.klog - st.lsplit)) * max(1, tiles)
    if isinstance(st, JoinDesc):
        return (1 << st.s) * max(1, (st.nbatch << max(0, st.fx + st.fy - 8)) // max(1, st.nblocks))
    return 0


def compile_program(tree, tables, num_groups, entry_type, W, device_alloc, regroup=True, group_levels=True):
    """Build a Program for batch size W.

    ``device_alloc(nbytes)`` returns an object with ``.data_ptr()`` (device memory that
    stays alive as long as the Program).  ``tables`` maps tensor id -> LeafTable.  With
    ``group_levels`` the joins run level by level as grouped launches (one per kernel
    variant and level) when that keeps the arena within ``LEVEL_PEAK_FACTOR`` of the
    post-order peak; otherwise, and for per-join timing, one launch per join."""
    esize = 8 if entry_type == _lib.TO_F64 else 4
    scaled = entry_type == _lib.TO_F32S
    fdtype = np.float64 if esize == 8 else np.float32
    nodes = analyse(tree, tables, regroup)
    prog = Program()
    prog.nbatch = W
    prog.nodes = nodes

    for n in nodes:
        if not n.is_leaf:
            nb = W if n.batched else 1
            g = join_geometry(nodes[n.left].rank, nodes[n.right].rank, n.lmap, n.rmap)
            g["var"] = choose_variant(g, nb, esize)
            n.geom = g
    order = choose_order(nodes, tables, W, esize)[0] if group_levels else "post"
    peak, partial = plan_memory(nodes, tables, W, esize, order)
    prog.arena_bytes = peak
    prog.order = order

    # ---- leaf tables (host) -> one device buffer; the scaled executor's leaf factors
    # multiply the count (prog.leaf_factor, mantissa and exponent)
    leaf_ids = sorted({n.tensor for n in nodes if n.is_leaf})
    table_off, chunks, exp0, factor_of, builds = {}, [], {}, {}, []
    off = 0
    for t in leaf_ids:
        tab = tables[t]
        if tab.spec is not None:  # written on the device by the leaf build launch
            v0, v1, F, e = tab.spec.stored(scaled)
            builds.append((t, v0, v1))
            size = tab.size
        else:
            rows = tab.rows
            if scaled:
                rows, F, e = leaf_encoding(rows)
            arr = np.ascontiguousarray(rows, dtype=fdtype).ravel()
            chunks.append((t, arr))
            size = arr.size
        if scaled:
            exp0[t] = e
            factor_of[t] = F
        table_off[t] = off
        off += (size * esize + ALIGN - 1) // ALIGN * ALIGN // esize
    table_host = np.zeros(max(off, 1), dtype=fdtype)
    for t, arr in chunks:
        table_host[table_off[t]:table_off[t] + arr.size] = arr
    # one factor per leaf *occurrence*: a tree may hold one tensor at several leaves (the
    # line planner joins a rank-0 tensor at its first bag and again at the top,
    # line_graph_method.py:16-35 and contraction_tree.pyx:302-310), and each occurrence
    # multiplies the count by that tensor's values
    prog.leaf_factor = factor_product([factor_of[n.tensor] for n in nodes if n.is_leaf] if scaled else [])
    coef_host = np.zeros(max(1, len(leaf_ids) * max(1, num_groups)), dtype=np.int64)
    coef_off = {}
    for i, t in enumerate(leaf_ids):
        c = tables[t].coef
        coef_off[t] = i * max(1, num_groups)
        coef_host[coef_off[t]:coef_off[t] + len(c)] = c

    # ---- scale records: [constant leaves][constant joins][wave tensors]
    nrec = 0
    rec_init = []
    if scaled:
        for n in nodes:
            if n.is_leaf and not n.batched:
                n.rec, n.rec_stride = nrec, 0
                rec_init.append((nrec, exp0[n.tensor]))
                nrec += 1
        pro_rec0 = nrec
        for n in nodes:
            if not n.is_leaf and not n.batched:
                n.rec, n.rec_stride = nrec, 0
                nrec += 1
        wave_rec0 = nrec
        for n in nodes:
            if n.batched:
                n.rec, n.rec_stride = nrec, 1
                nrec += W
        prog.rec_ranges = (pro_rec0, wave_rec0, nrec)

    # ---- device buffers
    prog.table = device_alloc(table_host.nbytes)
    prog.coef = device_alloc(coef_host.nbytes)
    prog.arena = device_alloc(max(ALIGN, peak))
    prog.recs = device_alloc(max(1, nrec) * ctypes.sizeof(_lib.ScaleRec))
    prog.params = device_alloc(4 * 8)  # wave params [begin, end], prologue params [0, 1]
    prog.accum = device_alloc(ctypes.sizeof(_lib.Accum))
    prog.table_host, prog.coef_host = table_host, coef_host
    tbase, abase, rbase = prog.table.data_ptr(), prog.arena.data_ptr(), prog.recs.data_ptr()
    rec_size = ctypes.sizeof(_lib.ScaleRec)
    prog.rec_init = rec_init
    wave_params = prog.params.data_ptr()
    pro_params = wave_params + 16

    for n in nodes:
        if n.is_leaf and not n.batched:
            n.ptr = tbase + table_off[n.tensor] * esize
            n.bstride = 0
        else:
            n.ptr = abase + n.buf[0]
            n.bstride = (1 << n.rank) if n.batched else 0

    def recp(n):
        return rbase + n.rec * rec_size if scaled else 0

    # ---- descriptors
    blob = bytearray()

    def put(structs):
        nonlocal blob
        start = (len(blob) + ALIGN - 1) // ALIGN * ALIGN
        blob += b"\0" * (start - len(blob))
        for st in structs:
            blob += bytes(st)
        return start

    pending = []  # (list, Launch, blob offset attr)

    # leaf tables written on the device once, right after upload
    prog.build = []
    if builds:
        bdescs, block0 = [], 0
        for t, v0, v1 in builds:
            bd = tables[t].spec.desc(tbase + table_off[t] * esize, v0, v1)
            bd.block0 = block0
            block0 += bd.nblocks
            bdescs.append(bd)
        boff = put(bdescs)
        prog.build.append((Launch(op=_lib.TO_OP_LEAF_BUILD, count=len(bdescs), grid=block0), boff))

    def join_launches(n, phase_list):
        for kind, op, variant, lds, st, nblk in join_records(n, nodes, W, esize, scaled, abase,
                                                               partial.get(n.idx, {}), recp):
            off = put([st])
            if kind == "permute":
                work = (2 * esize * (int(st.nbatch) << int(st.rank)), 0)
            elif kind in ("stream", "splitk", "gemm"):
                work = join_traffic(n, nodes, W, esize)
            else:
                work = (0, 0)
            phase_list.append((kind, Launch(op=op, variant=variant, count=1, grid=nblk, lds=lds), off, n.idx,
                               work))

    def level_launches(level, phase_list):
        # every join of one level, grouped by kernel variant; permutes before products,
        # products before split-K finishing
        groups = {}
        for i in level:
            n = nodes[i]
            for kind, op, variant, lds, st, nblk in join_records(n, nodes, W, esize, scaled, abase,
                                                                   partial.get(n.idx, {}), recp):
                if kind == "permute":
                    work = (2 * esize * (int(st.nbatch) << int(st.rank)), 0)
                elif kind in ("stream", "splitk", "gemm"):
                    work = join_traffic(n, nodes, W, esize)
                else:
                    work = (0, 0)
                # small streamed joins of one level share one launch whatever their form
                # (each descriptor names its own): they are bound by launch latency
                if kind == "stream" and work[0] < STREAM_MERGE_BYTES:
                    variant = _lib.TO_STREAM_ANY
                key = (KIND_ORDER[kind], op, variant)
                groups.setdefault(key, []).append((kind, st, nblk, lds, n.idx, work))
        for key in sorted(groups):
            # the joins whose blocks run longest start first (blocks start in block order)
            grp = sorted(groups[key], key=lambda g: -block_cost(g[1]))
            block0, lds, wb, wf, structs = 0, 0, 0, 0, []
            for kind, st, nblk, l, idx, work in grp:
                st.block0 = block0
                block0 += nblk
                lds = max(lds, l)
                wb += work[0]
                wf += work[1]
                structs.append(st)
            off = put(structs)
            phase_list.append((grp[0][0], Launch(op=key[1], variant=key[2], count=len(structs), grid=block0,
                                                 lds=lds), off, grp[0][4] if len(grp) == 1 else None, (wb, wf)))

    pro, wave = [], []
    if scaled and prog.rec_ranges[1] > prog.rec_ranges[0]:
        a, b = prog.rec_ranges[0], prog.rec_ranges[1]
        pro.append(("clear", Launch(op=_lib.TO_OP_SCALE_CLEAR, grid=max(1, min(1024, -(-(b - a) // 256))),
                                    aux0=rbase + a * rec_size, iaux0=b - a), None, None, (0, 0)))
    if order == "level":
        for level in join_levels(nodes, False):
            level_launches(level, pro)
    else:
        for n in nodes:
            if not n.is_leaf and not n.batched:
                join_launches(n, pro)
    root = nodes[-1]
    prog.root_const = not root.batched
    if scaled and prog.rec_ranges[2] > prog.rec_ranges[1]:
        a, b = prog.rec_ranges[1], prog.rec_ranges[2]
        wave.append(("clear", Launch(op=_lib.TO_OP_SCALE_CLEAR, grid=max(1, min(1024, -(-(b - a) // 256))),
                                     aux0=rbase + a * rec_size, iaux0=b - a), None, None, (0, 0)))
    leaf_descs = []
    block0 = 0
    for n in nodes:
        if n.is_leaf and n.batched:
            ld = LeafDesc()
            ld.src = tbase + table_off[n.tensor] * esize
            ld.dst = n.ptr
            ld.coef = prog.coef.data_ptr() + coef_off[n.tensor] * 8
            ld.rec = recp(n)
            ld.exp0 = exp0.get(n.tensor, 0)
            ld.rank = n.rank
            ld.ngroups = num_groups
            ld.nbatch = W
            ld.block0 = block0
            ld.nblocks = int(min(1024, -(-(W << n.rank) // 256)))
            block0 += ld.nblocks
            leaf_descs.append(ld)
    if leaf_descs:
        off = put(leaf_descs)
        leaf_bytes = sum(esize * (W << ld.rank) for ld in leaf_descs)
        wave.append(("leaf", Launch(op=_lib.TO_OP_LEAF_GATHER, count=len(leaf_descs), grid=block0,
                                    aux3=wave_params), off, None, (leaf_bytes, 0)))
    if order == "level":
        for level in join_levels(nodes, True):
            level_launches(level, wave)
    else:
        for n in nodes:
            if not n.is_leaf and n.batched:
                join_launches(n, wave)
    root_launch = Launch(op=_lib.TO_OP_ROOT_REDUCE, grid=1, aux0=root.ptr,
                         aux1=recp(root) if scaled else 0, aux2=prog.accum.data_ptr(),
                         aux3=pro_params if prog.root_const else wave_params,
                         iaux0=1 if prog.root_const else W)
    if prog.root_const:
        pro.append(("root", root_launch, None, None, (0, 0)))
    else:
        wave.append(("root", root_launch, None, None, (0, 0)))
        wave.append(("advance", Launch(op=_lib.TO_OP_WAVE_ADVANCE, grid=1, aux3=wave_params, iaux0=W), None, None,
                     (0, 0)))

    prog.blob_host = np.frombuffer(bytes(blob) or b"\0", dtype=np.uint8).copy()
    prog.blob = device_alloc(prog.blob_host.nbytes)
    bbase = prog.blob.data_ptr()
    for L, boff in prog.build:
        L.descs = bbase + boff
    prog.build = [L for L, _ in prog.build]
    # per launch: (kind, node index) and its algorithmic (bytes, flops): each distinct
    # operand entry read once and each output entry written once
    prog.prologue_info, prog.wave_info = [], []
    prog.prologue_work, prog.wave_work = [], []
    for lst, out, info, work in ((pro, prog.prologue, prog.prologue_info, prog.prologue_work),
                                 (wave, prog.wave, prog.wave_info, prog.wave_work)):
        for kind, L, off, nidx, w in lst:
            if off is not None:
                L.descs = bbase + off
            out.append(L)
            info.append((kind, nidx))
            work.append(w)
    prog.stats = dict(
        nodes=len(nodes),
        joins=sum(1 for n in nodes if not n.is_leaf),
        batched_joins=sum(1 for n in nodes if not n.is_leaf and n.batched),
        splitk=sum(1 for n in nodes if not n.is_leaf and n.geom["var"]["kind"] == "splitk"),
        gemm=sum(1 for n in nodes if not n.is_leaf and n.geom["var"]["kind"] == "gemm"),
        arena_bytes=peak,
        table_bytes=int(table_host.nbytes),
        device_built_leaves=len(builds),
        launches_prologue=len(prog.prologue),
        launches_wave=len(prog.wave),
    )
    return prog


def join_traffic(node, nodes, W, esize):
    """Algorithmic (bytes, flops) of one join: each distinct operand entry read once,
    each output entry written once; flops = 2 * outputs * 2^s."""
    nb = W if node.batched else 1
    g = node.geom
    total = 0
    for ch in (nodes[node.left], nodes[node.right]):
        total += (1 << ch.rank) * (nb if ch.batched else 1)
    total += (1 << node.rank) * nb
    flops = 2 * nb * (1 << (g["fx"] + g["fy"] + g["s"]))
    return total * esize, flops


def rec_init_bytes(prog):
    """Initial scale records (constant leaves: exponent exp0, max |entry| 1.0)."""
    n = max(1, prog.rec_ranges[2] if hasattr(prog, "rec_ranges") else 1)
    recs = (_lib.ScaleRec * n)()
    one = np.array([1.0], dtype=np.float32).view(np.uint32)[0]
    for i, e in getattr(prog, "rec_init", []):
        recs[i].exp = e
        recs[i].maxbits = int(one)
    return bytes(recs)
