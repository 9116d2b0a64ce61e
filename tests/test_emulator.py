"""The schedule compiler's launch records, executed by the test-only numpy emulator,
reproduce the oracle: bit placement, batching, slice-row gather, memory plan, split-K and
the (mantissa, exponent) slice sum, in both entry types."""

import ctypes
import json
import os

import numpy as np
import pytest
from fixtures import plan_fixture_paths

import emulator
from oracle import numpy_oracle
from tensororder_amd import _lib, plan_io, schedule
from tensororder_amd.network import BuiltTensor, TensorNetwork
from tensororder_amd.plan import SlicedExecutionPlan
from tensororder_amd.tree import TreeContext

HERE = os.path.dirname(os.path.abspath(__file__))
SMALL = [f for f in plan_fixture_paths() if json.load(open(f))["meta"]["maxrank"] <= 16]


@pytest.mark.parametrize("path", SMALL, ids=[os.path.basename(p)[:-5] for p in SMALL])
@pytest.mark.parametrize("et,tol", [(_lib.TO_F64, 1e-12), (_lib.TO_F32S, 1e-6)])
def test_fixture_emulated(path, et, tol):
    doc = json.load(open(path))
    plan = plan_io.plan_from_dict(doc)
    want = numpy_oracle.contract_sliced(doc)
    tree, tables, G, total = emulator.plan_tables(plan, emulator.HostOps)
    for W in sorted({1, 3, total}):
        got, prog = emulator.emulate_plan(tree, tables, G, total, et, W)
        assert got == pytest.approx(want, rel=tol)


def _two_tensor_plan(ra, rb, shared, seed):
    """A network of two dense tensors sharing `shared` edges (order scrambled)."""
    rng = np.random.default_rng(seed)
    net = TensorNetwork()
    a = net.add_node(BuiltTensor(rng.standard_normal((2,) * ra)))
    b = net.add_node(BuiltTensor(rng.standard_normal((2,) * rb)))
    pa = rng.permutation(ra)[:shared]
    pb = rng.permutation(rb)[:shared]
    for i, j in zip(pa, pb):
        net.connect(a[i][0], a[i][1], b[j][0], b[j][1])
    ctx = TreeContext()
    # remaining free axes get connected to rank-1 dense tensors so the root is a scalar
    leaves = []
    for t, r in ((0, ra), (1, rb)):
        for ax in range(r):
            if net._index_lists[t][ax] < 0:
                u = net.add_node(BuiltTensor(rng.standard_normal((2,))))
                net.connect(t, ax, u[0][0], 0)
                leaves.append(u[0][0])
    root = ctx.join(ctx.leaf(net, 0), ctx.leaf(net, 1))
    for u in leaves:
        root = ctx.join(root, ctx.leaf(net, u))
    return SlicedExecutionPlan(ctx.get_tree(root), net)


DENSE = [(12, 12, 12), (13, 11, 10), (9, 14, 9), (6, 6, 0), (10, 3, 1), (12, 12, 6), (14, 11, 5), (11, 15, 5),
         (10, 10, 4), (15, 15, 15), (16, 15, 14), (18, 17, 12)]


@pytest.mark.parametrize("ra,rb,s", DENSE)
@pytest.mark.parametrize("et,tol", [(_lib.TO_F64, 1e-11), (_lib.TO_F32S, 2e-5)])
def test_dense_joins_emulated(ra, rb, s, et, tol):
    plan = _two_tensor_plan(ra, rb, s, seed=ra * 100 + rb)
    doc = plan_io.plan_to_dict(plan)
    want = numpy_oracle.contract_sliced(doc)
    tree, tables, G, total = emulator.plan_tables(plan, emulator.HostOps)
    got, prog = emulator.emulate_plan(tree, tables, G, total, et, 1)
    if s >= 9 and ra + rb - 2 * s <= 6:
        assert prog.stats["splitk"] >= 1
    if s >= schedule.SPLITK_PERMUTE_MIN_SHARED:  # scrambled shared bits: operands permuted first
        kinds = [k for k, _ in prog.prologue_info + prog.wave_info]
        assert "splitk" in kinds and "permute" in kinds
    if ra - s >= 6 and rb - s >= 6 and s >= 4:
        assert prog.stats["gemm"] >= 1
    assert got == pytest.approx(want, rel=tol, abs=tol * 1e-3 * abs(want))


def test_geometry_is_numpy_order():
    # output index = (left free index << b) | right free index
    g = schedule.join_geometry(4, 3, [1, 3], [2, 0])
    assert g["fx"] + g["fy"] == 3 and g["s"] == 2
    rng = np.random.default_rng(0)
    a = rng.standard_normal((2,) * 4)
    b = rng.standard_normal((2,) * 3)
    want = np.tensordot(a, b, ([1, 3], [2, 0])).ravel()
    A, B = a.ravel(), b.ravel()
    X, Y = (A, B) if g["x_is_left"] else (B, A)
    out = np.zeros(8)
    for f in range(1 << g["fx"]):
        for y in range(1 << g["fy"]):
            acc = 0.0
            for k in range(1 << g["s"]):
                xo = emulator.pdep_array(np.array([f]), g["xf"], g["fx"])[0] + \
                    emulator.pdep_array(np.array([k]), g["xs"], g["s"])[0]
                yo = emulator.pdep_array(np.array([y]), g["yf"], g["fy"])[0] + \
                    emulator.pdep_array(np.array([k]), g["ys"], g["s"])[0]
                acc += X[xo] * Y[yo]
            out[(f << g["shf"]) | (y << g["shy"])] = acc
    np.testing.assert_allclose(out, want, rtol=1e-13)


@pytest.mark.parametrize("order", ["post", "level"])
def test_memory_plan_never_overlaps_live_buffers(order):
    # replay the execution groups (one join, or every join of a level): the outputs and
    # scratch buffers of a group must not overlap each other nor any buffer that is still
    # needed (an operand of this group or of a later one)
    doc = json.load(open(os.path.join(HERE, "golden", "cnf40_factor_s6.json")))
    plan = plan_io.plan_from_dict(doc)
    tree, tables, G, total = emulator.plan_tables(plan, emulator.HostOps)
    nodes = schedule.analyse(tree, tables)
    consumer = {}
    for n in nodes:
        if not n.is_leaf:
            for c in (n.left, n.right):
                consumer[c] = n.idx
    for W in (1, 5, 64):
        schedule.peak_bytes(nodes, tables, W, _lib.TO_F64)
        peak, scratch = schedule.plan_memory(nodes, tables, W, 8, order)
        if order == "level":
            groups = schedule.join_levels(nodes, False) + [None] + schedule.join_levels(nodes, True)
        else:
            groups = [[n.idx] for n in nodes if not n.is_leaf and not n.batched] + [None] + \
                     [[n.idx] for n in nodes if not n.is_leaf and n.batched]
        done = set()
        for grp in groups:
            if grp is None:  # wave start: batched leaves are gathered
                done |= {n.idx for n in nodes if n.is_leaf and n.batched}
                continue
            fresh = [nodes[i].buf for i in grp] + [blk for i in grp for blk in scratch.get(i, {}).values()]
            for k, (s0, z0) in enumerate(fresh):
                assert s0 + z0 <= peak
                for s1, z1 in fresh[k + 1:]:
                    assert s0 + z0 <= s1 or s1 + z1 <= s0, ("group overlap", grp)
                for m in done:
                    if nodes[m].buf is None or consumer.get(m) is None or consumer[m] in done:
                        continue
                    s1, z1 = nodes[m].buf
                    assert s0 + z0 <= s1 or s1 + z1 <= s0, ("live overlap", grp, m)
            done |= set(grp)


@pytest.mark.parametrize("et,tol", [(_lib.TO_F64, 1e-12), (_lib.TO_F32S, 1e-6)])
def test_batch_of_uneven_instances_emulated(et, tol, monkeypatch):
    from tensororder_amd import batch, network, planners
    from tensororder_amd.formula import parse_dimacs_text
    from tensororder_amd.instances import vertex_cover_cnf

    instances, wants = [], []
    for seed, n in enumerate([10, 16, 24, 12, 30, 20]):
        f = parse_dimacs_text(vertex_cover_cnf(n, seed, True), "cachet")
        plan = planners.best_plan(network.cnf_count(f), "line", solver=None)
        wants.append(numpy_oracle.contract_sliced(plan_io.plan_to_dict(plan)))
        tree, tables, G, total = emulator.plan_tables(plan, emulator.HostOps)
        instances.append((tree, tables))
    # with one-wave joins turned off, block -> descriptor maps on every grouped launch of
    # two or more descriptors, each checked by the emulator against the binary search the
    # kernels do without one
    with monkeypatch.context() as mp:
        mp.setattr(batch, "MAP_MIN_DESCS", 2)
        mp.setattr(schedule, "WAVE_MAX_LANE_PRODUCTS", 0)
        got, prog = emulator.emulate_batch(instances, et)
    mapped = sum(1 for L in prog.launches if L.aux1 and L.op in emulator.Emulator.TABLE_DESCS)
    assert mapped > 0 and prog.maps_checked == mapped and prog.wave_joins == 0
    for g, w in zip(got, wants):
        assert g == pytest.approx(w, rel=tol)
    # as compiled: the smallest streamed joins run one wave each
    got, prog = emulator.emulate_batch(instances, et)
    assert prog.wave_joins > 0
    assert prog.stats["steps"] < prog.stats["joins"]
    for g, w in zip(got, wants):
        assert g == pytest.approx(w, rel=tol)
    # graph branches: streamed and split-K joins (with their operand permutes) go to the
    # side stream, products stay on the main one, and every step starts with a join of the
    # two branches (so a step sees all of the previous one)
    first = {}
    side_ops = {_lib.TO_OP_JOIN_STREAM, _lib.TO_OP_JOIN_SPLITK, _lib.TO_OP_JOIN_FINISH, _lib.TO_OP_PERMUTE}
    for i, (L, (kind, k)) in enumerate(zip(prog.launches, prog.info)):
        if L.flags & _lib.TO_LAUNCH_SIDE:
            assert L.op in side_ops
        if L.op == _lib.TO_OP_JOIN_STREAM:
            assert L.flags & _lib.TO_LAUNCH_SIDE
        if k >= 0:
            first.setdefault(k, i)
    assert any(L.flags & _lib.TO_LAUNCH_SIDE for L in prog.launches)
    for k, i in first.items():
        assert prog.launches[i].flags & _lib.TO_LAUNCH_JOIN, k
    assert sum(1 for L in prog.launches if L.flags & _lib.TO_LAUNCH_JOIN) == len(first)


def _gemm_descs(prog):
    """(variant, GemmDesc) of every MFMA product launch of a compiled program."""
    out = []
    for L in list(prog.prologue) + list(prog.wave):
        if L.op != _lib.TO_OP_GEMM:
            continue
        base = L.descs - prog.blob.data_ptr()
        for i in range(L.count):
            off = base + i * ctypes.sizeof(_lib.GemmDesc)
            out.append((L.variant, _lib.GemmDesc.from_buffer_copy(prog.blob_host.tobytes()[off:off + ctypes.sizeof(_lib.GemmDesc)])))
    return out


@pytest.mark.parametrize("et,tol", [(_lib.TO_F64, 1e-11), (_lib.TO_F32S, 2e-5)])
def test_batch_split_k_products_emulated(et, tol, monkeypatch):
    # products of a batch step with few 64x64 tiles over a long K are split into K parts
    # whose partial sums a finishing launch adds (the emulator writes and sums the parts in
    # the documented layout); every count against the oracle.  Parts of >= 128 terms here,
    # so that products small enough to emulate are split
    monkeypatch.setattr(schedule, "GEMM_SPLIT_MIN_KLOG", 7)
    instances, wants = [], []
    for i, (ra, rb, s) in enumerate(((16, 16, 10), (15, 16, 9), (17, 16, 10), (14, 14, 8))):
        plan = _two_tensor_plan(ra, rb, s, seed=50 + i)
        wants.append(numpy_oracle.contract_sliced(plan_io.plan_to_dict(plan)))
        tree, tables, G, total = emulator.plan_tables(plan, emulator.HostOps)
        instances.append((tree, tables))
    got, prog = emulator.emulate_batch(instances, et)
    assert prog.split_gemms >= 3
    gemms = [L for L in prog.launches if L.op == _lib.TO_OP_GEMM]
    finishes = [L for L in prog.launches if L.op == _lib.TO_OP_JOIN_FINISH]
    assert gemms and finishes
    for L in gemms:
        assert L.variant in (0, 2)  # the 64x64 form, the only one that reads lsplit
    for g, w in zip(got, wants):
        assert g == pytest.approx(w, rel=tol, abs=tol * 1e-3 * abs(w))


@pytest.mark.parametrize("et", [_lib.TO_F64, _lib.TO_F32S])
def test_gemm_descriptors_fit_their_tiles(et):
    # every product the schedule hands to the MFMA kernels satisfies the tile shape the
    # variant assumes: rows/columns >= the tile edge, shared range >= one K slab
    cases = []
    for path in plan_fixture_paths():
        doc = json.load(open(path))
        if doc["meta"]["maxrank"] <= 20:
            cases.append(plan_io.plan_from_dict(doc))
    for ra, rb, s in ((13, 13, 5), (14, 13, 6), (16, 14, 4), (18, 17, 11), (12, 12, 6)):
        cases.append(_two_tensor_plan(ra, rb, s, seed=1))
    f32 = et == _lib.TO_F32S
    seen = tall = 0
    for plan in cases:
        tree, tables, G, total = emulator.plan_tables(plan, emulator.HostOps)
        heap = emulator.Heap()
        prog = schedule.compile_program(tree, tables, G, et, min(total, 4), heap.alloc)
        for variant, gd in _gemm_descs(prog):
            seen += 1
            if variant & 1:
                assert gd.mlog >= 7 and gd.nlog >= (6 if (variant & 2 and not variant & 8) else 7)
                assert gd.klog >= (5 if f32 else 4)  # K slab of 32 floats / 16 doubles
            else:
                assert gd.mlog >= 6 and gd.nlog >= 6 and gd.klog >= 4
            assert not (variant & 2) or f32
            if variant & 16:  # bf16 split products: float, 128x128 tiles, one block per CU
                assert f32 and variant & 1 and variant & 8 and gd.nlog >= 7
                assert gd.klog >= (7 if variant & 2 else 5)  # wide: whole partial sums of 4 slabs
            if variant & 32:  # 256-row split products: float accumulators only
                tall += 1
                assert variant & 16 and not variant & 2 and gd.mlog >= 8
            assert variant in (0, 1, 2, 3, 5, 7, 9, 11, 13, 15, 25, 27, 29, 31, 57, 61), variant
    assert seen > 0
    assert tall > 0 or not f32


# DIMACS inputs that declare variables no clause uses.  The line planner matches a rank-0
# variable tensor at its first bag and joins it again at the top (the reference's
# line_graph_method.py:16-35 with contraction_tree.pyx:302-310), so one tensor sits at two
# leaves and the count carries its weight sum twice; the oracle reproduces that, and the
# FP32-scaled executor's leaf factors must follow every leaf occurrence.
UNUSED_VARIABLE_CNFS = [
    ("unweighted", "p cnf 5 3\n1 -2 0\n2 3 0\n-1 -3 0\n"),
    ("cachet", "p cnf 6 3\nw 1 0.3\nw 2 0.7\nw 4 -1\nw 5 0.125\n1 -2 0\n2 3 0\n-1 -3 2 0\n"),
    ("mcc", "p cnf 5 2\nw 1 0.25\nw -1 0.5\nw 4 3.5\nw -4 0.75\nw 5 2\nw -5 2\n1 -2 0\n2 3 0\n"),
]


def unused_variable_plans():
    from tensororder_amd import network, planners
    from tensororder_amd.formula import parse_dimacs_text

    out = []
    for wf, text in UNUSED_VARIABLE_CNFS:
        f = parse_dimacs_text(text, wf)
        for planner in ("line", "factor"):
            out.append(("%s-%s" % (wf, planner), planners.best_plan(network.cnf_count(f), planner, solver=None)))
    return out


@pytest.mark.parametrize("et,tol", [(_lib.TO_F64, 1e-12), (_lib.TO_F32S, 1e-6)])
def test_unused_variables_emulated(et, tol):
    plans = unused_variable_plans()
    repeated = 0
    instances, wants = [], []
    for name, plan in plans:
        want = numpy_oracle.contract_sliced(plan_io.plan_to_dict(plan))
        tree, tables, G, total = emulator.plan_tables(plan, emulator.HostOps)
        leaves = [t.tensor_index for t in tree.iterate_postorder() if t.is_leaf]
        repeated += len(leaves) - len(set(leaves))
        got, _ = emulator.emulate_plan(tree, tables, G, total, et, 1)
        assert got == pytest.approx(want, rel=tol), name
        instances.append((tree, tables))
        wants.append(want)
    # the line plans hold repeated leaves (otherwise this test would not test them)
    assert repeated > 0
    got, _ = emulator.emulate_batch(instances, et)
    for (name, _), g, w in zip(plans, got, wants):
        assert g == pytest.approx(w, rel=tol), name


class _NumpyOps:
    def create_tensor(self, shape, default_value=None):
        return np.full(shape, 0.0 if default_value is None else default_value, dtype=np.float64)

    def tensordot(self, a, b, axes):
        return np.tensordot(a, b, axes)


def early_contracted_plan(early, nvars=26, nclauses=60, seed=5, slices=3):
    """A sliced factor plan after the reference's early contraction
    (sliced_execution_plan.py:29-56): rerooted, regrouped, joins below ``early`` already
    contracted into dense leaves."""
    from tensororder_amd import network, planners
    from tensororder_amd.formula import parse_dimacs_text
    from tensororder_amd.instances import random_3cnf
    from tensororder_amd.plan import ALL_SLICERS

    f = parse_dimacs_text(random_3cnf(nvars, nclauses, seed), "cachet")
    plan = planners.best_plan(network.cnf_count(f), "factor", solver=None, seed=0)
    ALL_SLICERS["greedy_mem"].slice_until(plan, slices=slices)
    want = numpy_oracle.contract_sliced(plan_io.plan_to_dict(plan))
    plan.contract_small(early, _NumpyOps())
    return plan, want


@pytest.mark.parametrize("early", [4, 6, 9])
@pytest.mark.parametrize("et,tol", [(_lib.TO_F64, 1e-12), (_lib.TO_F32S, 1e-6)])
def test_early_contracted_plan_emulated(early, et, tol):
    plan, want = early_contracted_plan(early)
    doc = plan_io.plan_to_dict(plan)
    assert any(t["kind"] == "dense" for t in doc["tensors"])
    assert numpy_oracle.contract_sliced(doc) == pytest.approx(want, rel=1e-12)
    tree, tables, G, total = emulator.plan_tables(plan, emulator.HostOps)
    for W in sorted({1, total}):
        got, _ = emulator.emulate_plan(tree, tables, G, total, et, W)
        assert got == pytest.approx(want, rel=tol)


@pytest.mark.parametrize("et,tol", [(_lib.TO_F64, 1e-12), (_lib.TO_F32S, 1e-6)])
def test_batch_admission_waves_emulated(et, tol):
    # a budget below the whole batch's arena: the instances run in waves, in order, through
    # one arena of the largest wave's size, and every count still matches the oracle
    from tensororder_amd import batch, network, planners
    from tensororder_amd.formula import parse_dimacs_text
    from tensororder_amd.instances import vertex_cover_cnf

    instances, wants = [], []
    for seed, n in enumerate([10, 16, 24, 12, 30, 20, 18, 26]):
        f = parse_dimacs_text(vertex_cover_cnf(n, seed, True), "cachet")
        plan = planners.best_plan(network.cnf_count(f), "line", solver=None)
        wants.append(numpy_oracle.contract_sliced(plan_io.plan_to_dict(plan)))
        tree, tables, G, total = emulator.plan_tables(plan, emulator.HostOps)
        instances.append((tree, tables))
    per = batch.analyse_batch(instances, et)
    need = [p["region"] + p["fixed"] for p in per]
    budget = max(need) + sum(need) // 3
    waves = batch.admission_waves(per, budget)
    assert len(waves) >= 2 and sorted(i for w in waves for i in w) == list(range(len(per)))
    assert all(sum(need[i] for i in w) <= budget for w in waves)
    assert batch.admission_waves(per, max(need) - 1) is None
    got, progs = emulator.emulate_batch(instances, et, budget=budget)
    assert len(progs) == len(waves)
    for g, w in zip(got, wants):
        assert g == pytest.approx(w, rel=tol)


@pytest.mark.parametrize("et", [_lib.TO_F64, _lib.TO_F32S])
def test_device_leaf_build_matches_host_tables(et):
    # every OR / variable leaf kind under several slicings: the table the leaf build op
    # writes (emulated from its documented semantics) equals the host's reordered table,
    # or its float32 encoding
    import leafcases

    from tensororder_amd.hip_api import leaf_spec_table

    scaled = et == _lib.TO_F32S
    dt = np.float32 if scaled else np.float64
    n = 0
    for name, t, lookup, ngroups, rows, coef in leafcases.cases():
        tab = leaf_spec_table(t, lookup, ngroups)
        assert tab is not None and tab.coef == list(coef) and tab.nrows == rows.shape[0], name
        heap = emulator.Heap(1 << 20)
        buf = heap.alloc(tab.size * 8)
        v0, v1, F, e = tab.spec.stored(scaled)
        bd = tab.spec.desc(buf.data_ptr(), v0, v1)
        blob = heap.alloc(ctypes.sizeof(bd))
        heap.write(blob, bytes(bd))
        emu = emulator.Emulator(heap, et)
        emu.run([_lib.Launch(op=_lib.TO_OP_LEAF_BUILD, count=1, grid=bd.nblocks, descs=blob.data_ptr())])
        got = heap.view(buf.data_ptr(), dt, tab.size).reshape(rows.shape)
        if scaled:
            want, F2, e2 = schedule.leaf_encoding(rows)
            assert (F, e) == (F2, e2), name
            np.testing.assert_array_equal(got, want.astype(np.float32), err_msg=name)
        else:
            np.testing.assert_array_equal(got, rows, err_msg=name)
        n += 1
    assert n > 100


def _geom(fx, fy, s, yinner=True):
    return dict(fx=fx, fy=fy, s=s, shf=fy if yinner else 0, shy=0 if yinner else fx,
                xf=list(range(s, s + fx)), xs=list(range(s)), yf=list(range(s, s + fy)), ys=list(range(s)),
                x_is_left=True)


@pytest.mark.parametrize("esize", [4, 8])
def test_batch_launch_shapes(esize):
    # the many-instance batch's launch-bound choices stay within what their kernels take:
    # one-wave joins only for s <= 6 and <= WAVE_MAX_LANE_PRODUCTS products per lane;
    # split-K products only on the 64x64 form, with parts of >= 2^GEMM_SPLIT_MIN_KLOG terms;
    # split-K joins with >= 2^SPLITK_BATCH_MIN_LOG_STEPS steps of 256 per split
    waves = splits = 0
    for fx in range(0, 15):
        for fy in range(0, 12):
            for s in range(0, 16):
                v = schedule.choose_variant(_geom(fx, fy, s), 1, esize, launch_bound=True)
                if v.get("wave"):
                    waves += 1
                    assert s <= 6 and (-(-(1 << (fx + fy)) // 64) << s) <= schedule.WAVE_MAX_LANE_PRODUCTS
                    assert schedule.stream_variant(v, None) == _lib.TO_STREAM_WAVE
                if v["kind"] == "gemm" and v.get("lsplit"):
                    splits += 1
                    assert not v["tile"] & 1 and s - v["lsplit"] >= schedule.GEMM_SPLIT_MIN_KLOG
                    tiles = 1 << (fx - 6 + fy - 6)
                    assert (tiles << (v["lsplit"] - 1)) < schedule.GEMM_SPLIT_BLOCKS
                    assert v["nout"] == 1 << (fx + fy) and v["nsplit"] == 1 << v["lsplit"]
                if v["kind"] == "splitk" and s >= 8:
                    assert v["nsplit"] <= 1 << max(0, s - 8 - schedule.SPLITK_BATCH_MIN_LOG_STEPS)
                # outside the batch neither form is chosen
                w = schedule.choose_variant(_geom(fx, fy, s), 1, esize)
                assert not w.get("wave") and not w.get("lsplit")
    assert waves > 0 and splits > 0


def test_block_cost_orders_longest_first():
    # a grouped launch's descriptors go longest block first, and the block ranges stay
    # contiguous and increasing in descriptor order (what the kernels' search assumes)
    short, long_ = _lib.JoinDesc(), _lib.JoinDesc()
    short.fx, short.fy, short.s, short.nbatch, short.nblocks = 9, 3, 1, 1, 8
    long_.fx, long_.fy, long_.s, long_.nbatch, long_.nblocks = 9, 3, 6, 1, 8
    assert schedule.block_cost(long_) > schedule.block_cost(short)
    g1, g2 = _lib.GemmDesc(), _lib.GemmDesc()
    g1.mlog, g1.nlog, g1.klog, g1.nbatch, g1.nblocks = 6, 6, 12, 1, 1
    g2.mlog, g2.nlog, g2.klog, g2.nbatch, g2.nblocks, g2.lsplit = 6, 6, 12, 1, 8, 3
    assert schedule.block_cost(g1) > schedule.block_cost(g2)
    assert schedule.block_cost(_lib.PermDesc()) == 0
