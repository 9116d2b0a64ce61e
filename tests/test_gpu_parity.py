"""GPU parity: the HIP executor against the oracle and the committed fixtures.

Tolerances: plain FP64 executor rel <= 1e-12 of the FP64 oracle (exact for unweighted
counts < 2^53); FP32-scaled executor rel <= 1e-6.
"""

import itertools
import json
import os
import sys

import numpy as np
import pytest
from fixtures import plan_fixture_paths

from oracle import numpy_oracle
from tensororder_amd import plan_io
from tensororder_amd.execution import run
from tensororder_amd.plan import ALL_SLICERS

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = plan_fixture_paths()
# every fixture with a stored full count (the vertex covers up to n = 200 at rank 28, the
# sliced 3-CNFs); the 4096-slice 3-CNF stores a slice prefix and has its own test
FULL = [f for f in FIXTURES if "float64" in json.load(open(f))["meta"]["expected"]]
F64_TOL = 1e-12
F32_TOL = 1e-6


def _name(path):
    return os.path.basename(path)[:-5]


def _say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hip():
    from tensororder_amd.hip_api import HipAPI

    return HipAPI()


@pytest.fixture(scope="module")
def hip32():
    from tensororder_amd.hip_api import HipF32API

    return HipF32API()


def _expected(doc):
    exp = doc["meta"]["expected"]
    if "float64" in exp:
        return exp["float64"]
    return float(numpy_oracle.contract_sliced(doc))


@pytest.mark.parametrize("path", FULL, ids=[_name(p) for p in FULL])
def test_fixture_fp64(hip, path):
    doc = json.load(open(path))
    _say("case", _name(path), "fp64")
    want = _expected(doc)
    got = hip.contract_sliced(plan_io.plan_from_dict(doc))
    assert got == pytest.approx(want, rel=F64_TOL)
    if "bigint" in doc["meta"]["expected"]:
        assert int(got) == int(doc["meta"]["expected"]["bigint"])


@pytest.mark.parametrize("path", FULL, ids=[_name(p) for p in FULL])
def test_fixture_fp32_scaled(hip32, path):
    doc = json.load(open(path))
    _say("case", _name(path), "fp32s")
    want = _expected(doc)
    got = hip32.contract_sliced(plan_io.plan_from_dict(doc))
    assert got == pytest.approx(want, rel=F32_TOL)


@pytest.mark.parametrize("W", [1, 3, 64])
def test_batch_sizes_and_prefix(hip, W):
    path = os.path.join(HERE, "golden", "cnf20_factor_s4.json")
    doc = json.load(open(path))
    _say("case batch", W)
    hip.add_argument("max_batch", W)
    try:
        plan = plan_io.plan_from_dict(doc)
        assert hip.contract_sliced(plan) == pytest.approx(_expected(doc), rel=F64_TOL)
        for k in (1, 5, 16):
            want = numpy_oracle.contract_sliced(doc, num_slice_limit=k)
            assert hip.contract_sliced(plan, num_slice_limit=k) == pytest.approx(want, rel=F64_TOL)
            assert hip.last_result.slices == k
    finally:
        hip.add_argument("max_batch", 1 << 16)


def test_graph_and_eager_agree(hip):
    doc = json.load(open(os.path.join(HERE, "golden", "cnf40_factor_s6.json")))
    _say("case graph vs eager")
    plan = plan_io.plan_from_dict(doc)
    hip.add_argument("max_batch", 8)
    try:
        hip.add_argument("graph", True)
        g = hip.contract_sliced(plan)
        hip.add_argument("graph", False)
        e = hip.contract_sliced(plan)
    finally:
        hip.add_argument("graph", True)
        hip.add_argument("max_batch", 1 << 16)
    assert g == e
    assert g == pytest.approx(_expected(doc), rel=F64_TOL)


@pytest.mark.parametrize("ra,rb,s", [(12, 12, 6), (14, 11, 5), (11, 15, 5), (10, 10, 4), (13, 11, 10),
                                     (12, 12, 12), (13, 13, 5), (14, 13, 6), (16, 14, 4), (18, 17, 11),
                                     (15, 15, 15), (16, 15, 14), (22, 22, 22), (18, 17, 12), (17, 16, 13)])
def test_dense_join_plans(hip, hip32, ra, rb, s):
    from test_emulator import _two_tensor_plan

    _say("case dense", ra, rb, s)
    plan = _two_tensor_plan(ra, rb, s, seed=ra * 100 + rb)
    want = numpy_oracle.contract_sliced(plan_io.plan_to_dict(plan))
    got = hip.contract_sliced(plan)
    assert got == pytest.approx(want, rel=1e-11)
    got32 = hip32.contract_sliced(plan)
    assert got32 == pytest.approx(want, rel=2e-5)


def test_tensordot_gemm_shapes(hip):
    rng = np.random.default_rng(5)
    _say("case tensordot gemm")
    for ra, rb, s in ((13, 12, 6), (12, 14, 7), (10, 10, 4)):
        a = rng.standard_normal((2,) * ra)
        b = rng.standard_normal((2,) * rb)
        la = [int(x) for x in rng.permutation(ra)[:s]]
        lb = [int(x) for x in rng.permutation(rb)[:s]]
        got = hip.tensordot(a, b, (la, lb)).numpy()
        np.testing.assert_allclose(got, np.tensordot(a, b, (la, lb)), rtol=1e-11, atol=1e-11)


def test_tensordot_matches_numpy(hip):
    rng = np.random.default_rng(1)
    _say("case tensordot")
    for trial in range(40):
        ra, rb = int(rng.integers(0, 9)), int(rng.integers(0, 9))
        s = int(rng.integers(0, min(ra, rb) + 1))
        la = [int(x) for x in rng.permutation(ra)[:s]]
        lb = [int(x) for x in rng.permutation(rb)[:s]]
        a = rng.standard_normal((2,) * ra)
        b = rng.standard_normal((2,) * rb)
        got = hip.tensordot(a, b, (la, lb)).numpy()
        np.testing.assert_allclose(got, np.tensordot(a, b, (la, lb)), rtol=1e-12, atol=1e-12)


def test_tensordot_large_shared_splitk(hip):
    # few outputs, many shared bits: the split-K kernels
    rng = np.random.default_rng(2)
    _say("case splitk")
    a = rng.standard_normal((2,) * 14)
    b = rng.standard_normal((2,) * 13)
    la = list(range(1, 14))
    lb = list(rng.permutation(13))
    got = hip.tensordot(a, b, (la, [int(x) for x in lb])).numpy()
    want = np.tensordot(a, b, (la, [int(x) for x in lb]))
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-11)


def test_permute_matches_numpy(hip, hip32):
    # ranks below and above the tile sizes: scalar tiles (<= 12 bits) and the 16-byte
    # chunk kernel (tiles of 13 bits for float, 12 for double)
    rng = np.random.default_rng(3)
    _say("case permute")
    for api, dt in ((hip, np.float64), (hip32, np.float32)):
        for r in (0, 1, 3, 7, 12, 13, 14, 16, 18):
            a = rng.standard_normal((2,) * r).astype(dt)
            for perm in ([int(x) for x in rng.permutation(r)], list(range(r))[::-1]):
                got = np.asarray(api.permute(a, perm), dtype=dt)
                np.testing.assert_array_equal(got, np.transpose(a, perm))


def test_execution_run_slices_on_oom(hip):
    # a device budget too small for the unsliced plan: the library raises
    # OutOfMemoryError, execution.run slices once more and retries every slice
    doc = json.load(open(os.path.join(HERE, "golden", "vc90_weighted_line.json")))
    _say("case oom retry")
    plan = plan_io.plan_from_dict(doc)
    assert len(plan.groups_to_slice) == 0
    hip.add_argument("mem_limit", 80 * 1024)
    try:
        got = run(plan, hip, ALL_SLICERS["greedy_mem"], None)
    finally:
        hip.add_argument("mem_limit", None)
    assert len(plan.groups_to_slice) > 0
    assert got == pytest.approx(_expected(doc), rel=F64_TOL)


def test_batch_of_uneven_instances(hip, hip32):
    from tensororder_amd import network, planners
    from tensororder_amd.formula import parse_dimacs_text
    from tensororder_amd.instances import mixed_batch_sizes, vertex_cover_cnf

    _say("case batch of instances")
    plans, wants = [], []
    for seed, n in enumerate(mixed_batch_sizes(24, 7, low=20, high=60)):
        f = parse_dimacs_text(vertex_cover_cnf(n, seed, True), "cachet")
        plan = planners.best_plan(network.cnf_count(f), "line", solver=None)
        plans.append(plan)
        wants.append(numpy_oracle.contract_sliced(plan_io.plan_to_dict(plan)))
    for api, tol in ((hip, F64_TOL), (hip32, F32_TOL)):
        exe = api.prepare_batch(plans)
        try:
            for rep in range(3):  # graph replays, each checked
                res, lat = exe.run()
                assert [r.value for r in res] == pytest.approx(wants, rel=tol)
                assert all(l >= 0 for l in lat)
        finally:
            exe.close()


def test_batch_fixture_instances(hip, hip32):
    # all 512 instances of the mixed batch (FlowCutter line plans of max rank 6-22) against
    # the oracle counts stored with the fixture
    doc = json.load(open(os.path.join(HERE, "golden", "vc_batch512.json")))
    plans, wants = plan_io.batch_plans(doc)
    _say("case batch fixture (512 instances)")
    for api, tol in ((hip, F64_TOL), (hip32, F32_TOL)):
        exe = api.prepare_batch(plans)
        try:
            res, lat = exe.run()
            assert [r.value for r in res] == pytest.approx(wants, rel=tol)
        finally:
            exe.close()


def _gemm_variants(esize, mlog, nlog, klog):
    # K slabs: 16 doubles, 32 floats
    out = [0] + ([2] if esize == 4 else [])
    if mlog >= 7 and klog >= (5 if esize == 4 else 4) and nlog >= 7:
        out += [1, 5, 9, 13]
    if esize == 4 and mlog >= 7 and klog >= 5:
        if nlog >= 6:
            out += [3, 7]
        if nlog >= 7:
            out += [11, 15]
    if esize == 4 and mlog >= 7 and nlog >= 7 and klog >= 5:
        out += [25, 29]  # bf16 three-way split products
        if mlog >= 8:
            out += [57, 61]  # 256-row tiles (float accumulators)
        if klog >= 7:
            out += [27, 31]  # wide: K in whole partial sums of 4 slabs
    return out


@pytest.mark.parametrize("et", ["f64", "f32s"])
def test_one_wave_stream_joins(hip, hip32, et):
    # TO_STREAM_WAVE: small joins of several shapes in one launch (one wave each, the last
    # block partly filled), batched and not, scattered shared bits and both output layouts,
    # against float64 sums; FP32-scaled outputs carry their exponent and max in the record
    import ctypes

    import torch

    from tensororder_amd import _lib

    api = hip if et == "f64" else hip32
    esize = api.get_entry_size()
    dt = np.float64 if esize == 8 else np.float32
    rng = np.random.default_rng(5)
    dev = torch.device("cuda", 0)
    # (fx, fy, s, nbatch, y bits lowest in the output)
    geoms = [(2, 1, 1, 1, True), (3, 2, 0, 1, False), (0, 1, 3, 3, True), (5, 1, 2, 2, False),
             (4, 0, 2, 1, True), (9, 1, 1, 1, True), (6, 0, 0, 1, True)]

    def pdep(v, pos):
        return sum(((v >> i) & 1) << p for i, p in enumerate(pos))

    descs, keep, cases = [], [], []
    rsz = ctypes.sizeof(_lib.ScaleRec)
    for fx, fy, s, nb, yinner in geoms:
        rx, ry = fx + s, fy + s
        px, py = [int(v) for v in rng.permutation(rx)], [int(v) for v in rng.permutation(ry)]
        xf, xs = sorted(px[:fx]), px[fx:]
        yf, ys = sorted(py[:fy]), py[fy:]
        X = rng.uniform(-1.0, 2.0, (nb, 1 << rx)).astype(dt)
        Y = rng.uniform(-1.0, 2.0, (nb, 1 << ry)).astype(dt)
        tx, ty = torch.from_numpy(X).to(dev), torch.from_numpy(Y).to(dev)
        tc = torch.full((nb, 1 << (fx + fy)), float("nan"), dtype=tx.dtype, device=dev)
        d = _lib.JoinDesc()
        d.x, d.y, d.c = tx.data_ptr(), ty.data_ptr(), tc.data_ptr()
        d.xb, d.yb, d.cb = 1 << rx, 1 << ry, 1 << (fx + fy)
        d.fx, d.fy, d.s, d.nbatch = fx, fy, s, nb
        d.shf, d.shy = (fy, 0) if yinner else (0, fx)
        d.kvar = _lib.TO_STREAM_WAVE
        for arr, vals in ((d.xf_pos, xf), (d.xs_pos, xs), (d.yf_pos, yf), (d.ys_pos, ys)):
            for i, p in enumerate(vals):
                arr[i] = p
        recs = None
        if esize == 4:
            rs = np.zeros((3 * nb, 2), dtype=np.int64)
            for b in range(nb):
                rs[b] = (3, int(np.float32(np.abs(X[b]).max()).view(np.uint32)))
                rs[nb + b] = (-2, int(np.float32(np.abs(Y[b]).max()).view(np.uint32)))
            recs = torch.from_numpy(rs).to(dev)
            d.xs, d.ys, d.cs = recs.data_ptr(), recs.data_ptr() + nb * rsz, recs.data_ptr() + 2 * nb * rsz
            d.xsb = d.ysb = 1
        want = np.zeros((nb, 1 << (fx + fy)))
        for b in range(nb):
            for f in range(1 << fx):
                for y in range(1 << fy):
                    want[b, (f << d.shf) | (y << d.shy)] = sum(
                        float(X[b, pdep(f, xf) + pdep(k, xs)]) * float(Y[b, pdep(y, yf) + pdep(k, ys)])
                        for k in range(1 << s))
        descs.append(d)
        keep += [tx, ty, tc, recs]
        cases.append((tc, recs, want, nb))
    blob = torch.frombuffer(bytearray(b"".join(bytes(d) for d in descs)), dtype=torch.uint8).to(dev)
    torch.cuda.synchronize()
    _say("case one-wave joins", et, len(descs))
    L = _lib.Launch(op=_lib.TO_OP_JOIN_STREAM, variant=_lib.TO_STREAM_WAVE, count=len(descs),
                    grid=-(-len(descs) // 4), descs=blob.data_ptr())
    arr = _lib.launch_array([L])
    _lib.check(api._lib.to_run(ctypes.addressof(arr), 1, api._entry_type,
                               ctypes.c_void_p(api._stream.cuda_stream)), "to_run")
    api._stream.synchronize()
    for (tc, recs, want, nb), geom in zip(cases, geoms):
        c = tc.cpu().numpy().astype(np.float64)
        if esize == 8:
            np.testing.assert_allclose(c, want, rtol=1e-12, atol=1e-300, err_msg=str(geom))
            continue
        rs = recs.cpu().numpy()
        for b in range(nb):
            # represented value: stored * 2^exp of the output record; operands carry 2^3, 2^-2
            cexp = int(rs[2 * nb + b, 0])
            np.testing.assert_allclose(np.ldexp(c[b], cexp), np.ldexp(want[b], 1), rtol=2e-6, atol=1e-30,
                                       err_msg=str(geom))
            got_max = int(rs[2 * nb + b, 1]) & 0xFFFFFFFF
            assert got_max == int(np.float32(np.abs(c[b]).max()).view(np.uint32)), geom


@pytest.mark.parametrize("et", ["f64", "f32s"])
def test_split_k_product_and_finish(hip, hip32, et):
    # the 64x64 product with K split into 2^lsplit parts (double partial sums in the
    # documented layout) and the finishing launch that adds them, for both output layouts,
    # against the float64 product; FP32-scaled outputs carry their exponent and max
    import ctypes

    import torch

    from tensororder_amd import _lib

    api = hip if et == "f64" else hip32
    esize = api.get_entry_size()
    dt = np.float64 if esize == 8 else np.float32
    rng = np.random.default_rng(13)
    dev = torch.device("cuda", 0)
    rsz = ctypes.sizeof(_lib.ScaleRec)
    for mlog, nlog, klog, nb, lsplit in ((6, 6, 10, 1, 3), (7, 6, 9, 2, 2), (6, 7, 12, 1, 5)):
        for variant in ((0, 2) if esize == 4 else (0,)):
            for cols_contiguous in (True, False):
                _say("case split-K", et, mlog, nlog, klog, nb, lsplit, variant, cols_contiguous)
                M, N, K, S = 1 << mlog, 1 << nlog, 1 << klog, 1 << lsplit
                a = rng.uniform(0.5, 2.0, (nb, M, K)).astype(dt)
                b = rng.uniform(0.5, 2.0, (nb, N, K)).astype(dt)
                want = np.einsum("bmk,bnk->bmn", a.astype(np.float64), b.astype(np.float64))
                ta, tb = torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev)
                tc = torch.full((nb, M * N), float("nan"), dtype=ta.dtype, device=dev)
                tp = torch.full((nb * M * N * S,), float("nan"), dtype=torch.float64, device=dev)
                recs = None
                if esize == 4:
                    rs = np.zeros((3 * nb, 2), dtype=np.int64)
                    for i in range(nb):
                        rs[i] = (2, int(np.float32(np.abs(a[i]).max()).view(np.uint32)))
                        rs[nb + i] = (-1, int(np.float32(np.abs(b[i]).max()).view(np.uint32)))
                    recs = torch.from_numpy(rs).to(dev)
                gd = _lib.GemmDesc()
                gd.a, gd.b, gd.c = ta.data_ptr(), tb.data_ptr(), tp.data_ptr()
                gd.ab, gd.bb, gd.cb = M * K, N * K, M * N
                gd.ldcm, gd.ldcn = (N, 1) if cols_contiguous else (1, M)
                gd.mlog, gd.nlog, gd.klog = mlog, nlog, klog
                gd.nbatch, gd.block0, gd.nblocks, gd.lsplit = nb, 0, 5, lsplit
                fd = _lib.JoinDesc()
                fd.c, fd.cb, fd.partial, fd.nsplit = tc.data_ptr(), M * N, tp.data_ptr(), S
                fd.fx, fd.fy, fd.nbatch, fd.nblocks = mlog, nlog, nb, 3
                fd.shf, fd.shy = (nlog, 0) if cols_contiguous else (0, mlog)
                if recs is not None:
                    gd.as_, gd.bs, gd.cs = recs.data_ptr(), recs.data_ptr() + nb * rsz, recs.data_ptr() + 2 * nb * rsz
                    gd.asb = gd.bsb = 1
                    fd.xs, fd.ys, fd.cs = gd.as_, gd.bs, gd.cs
                    fd.xsb = fd.ysb = 1
                blob = torch.frombuffer(bytearray(bytes(gd) + bytes(fd)), dtype=torch.uint8).to(dev)
                torch.cuda.synchronize()
                launches = [_lib.Launch(op=_lib.TO_OP_GEMM, variant=variant, count=1, grid=gd.nblocks,
                                        descs=blob.data_ptr()),
                            _lib.Launch(op=_lib.TO_OP_JOIN_FINISH, variant=0, count=1, grid=fd.nblocks,
                                        descs=blob.data_ptr() + ctypes.sizeof(gd))]
                arr = _lib.launch_array(launches)
                _lib.check(api._lib.to_run(ctypes.addressof(arr), 2, api._entry_type,
                                           ctypes.c_void_p(api._stream.cuda_stream)), "to_run")
                api._stream.synchronize()
                c = tc.cpu().numpy().astype(np.float64)
                c = c.reshape(nb, M, N) if cols_contiguous else c.reshape(nb, N, M).transpose(0, 2, 1)
                if esize == 8:
                    np.testing.assert_allclose(c, want, rtol=1e-12)
                    continue
                r = recs.cpu().numpy()
                for i in range(nb):
                    # stored * 2^exp of the output record; the operands carry 2^2 and 2^-1
                    np.testing.assert_allclose(np.ldexp(c[i], int(r[2 * nb + i, 0])), np.ldexp(want[i], 1),
                                               rtol=2e-6)
                    assert int(r[2 * nb + i, 1]) & 0xFFFFFFFF == int(np.float32(np.abs(c[i]).max()).view(np.uint32))


@pytest.mark.parametrize("et", ["f64", "f32s"])
def test_gemm_every_tile_variant(hip, hip32, et):
    # every MFMA product kernel (64x64, 128x128, 128x64 wide, 128x128 one-block-per-CU
    # double-buffered, bf16 three-way split; plain and transposed) on the same operands,
    # both output layouts, with fewer blocks than tiles, against the float64 product
    import ctypes

    import torch

    from tensororder_amd import _lib

    api = hip if et == "f64" else hip32
    esize = api.get_entry_size()
    dt = np.float64 if esize == 8 else np.float32
    rng = np.random.default_rng(11)
    shapes = [(7, 7, 4), (8, 7, 6)] if esize == 8 else [(7, 7, 5), (7, 7, 6), (8, 8, 11), (8, 6, 12), (7, 8, 13),
                                                        (9, 7, 8)]
    dev = torch.device("cuda", 0)
    for mlog, nlog, klog in shapes:
        nb = 2
        M, N, K = 1 << mlog, 1 << nlog, 1 << klog
        a = rng.uniform(0.5, 2.0, (nb, M, K)).astype(dt)
        b = rng.uniform(0.5, 2.0, (nb, N, K)).astype(dt)
        # all-zero slabs (whole 32-deep K slabs of some tiles' rows or columns, and a whole
        # K range of one batch item), which the split products skip
        a[:, : M // 2, :32] = 0
        if K >= 64:
            b[:, N // 2:, 32:64] = 0
        if K >= 128:
            a[1, :, 64:128] = 0
            b[0, :, 96:128] = 0
        want = np.einsum("bmk,bnk->bmn", a.astype(np.float64), b.astype(np.float64))
        ta = torch.from_numpy(a).to(dev)
        tb = torch.from_numpy(b).to(dev)
        err = {}
        for variant in _gemm_variants(esize, mlog, nlog, klog):
            for cols_contiguous in (True, False):
                _say("case gemm", et, mlog, nlog, klog, "variant", variant, "ldcn==1" if cols_contiguous else "ldcm==1")
                tc = torch.full((nb, M * N), float("nan"), dtype=ta.dtype, device=dev)
                gd = _lib.GemmDesc()
                gd.a, gd.b, gd.c = ta.data_ptr(), tb.data_ptr(), tc.data_ptr()
                gd.ab, gd.bb, gd.cb = M * K, N * K, M * N
                gd.ldcm, gd.ldcn = (N, 1) if cols_contiguous else (1, M)
                recs = None
                if esize == 4:
                    rs = np.zeros((3 * nb, 2), dtype=np.int64)  # exp, maxbits | pad
                    for i in range(nb):
                        rs[i, 1] = int(np.float32(np.abs(a[i]).max()).view(np.uint32))
                        rs[nb + i, 1] = int(np.float32(np.abs(b[i]).max()).view(np.uint32))
                    recs = torch.from_numpy(rs).to(dev)
                    rsz = ctypes.sizeof(_lib.ScaleRec)
                    gd.as_, gd.bs, gd.cs = recs.data_ptr(), recs.data_ptr() + nb * rsz, recs.data_ptr() + 2 * nb * rsz
                    gd.asb = gd.bsb = 1
                gd.mlog, gd.nlog, gd.klog = mlog, nlog, klog
                rows = 64 if variant in (0, 2) else (256 if variant & 32 else 128)
                tiles = nb * (M // rows) * (N // (64 if variant in (0, 2, 3, 7) else 128))
                gd.nbatch, gd.block0, gd.nblocks = nb, 0, max(1, min(3, tiles))
                blob = torch.frombuffer(bytearray(bytes(gd)), dtype=torch.uint8).to(dev)
                torch.cuda.synchronize()
                L = _lib.Launch(op=_lib.TO_OP_GEMM, variant=variant, count=1, grid=gd.nblocks,
                                descs=blob.data_ptr())
                arr = _lib.launch_array([L])
                _lib.check(api._lib.to_run(ctypes.addressof(arr), 1, api._entry_type,
                                           ctypes.c_void_p(api._stream.cuda_stream)), "to_run")
                api._stream.synchronize()
                c = tc.cpu().numpy().reshape(nb, M, N) if cols_contiguous else \
                    tc.cpu().numpy().reshape(nb, N, M).transpose(0, 2, 1)
                if esize == 8:
                    np.testing.assert_allclose(c, want, rtol=1e-12)
                else:
                    out = recs.cpu().numpy()[2 * nb:]
                    for i in range(nb):
                        assert int(out[i, 1]) == int(np.float32(np.abs(c[i]).max()).view(np.uint32))
                        got = c[i].astype(np.float64) * 2.0 ** int(out[i, 0])
                        # float operands, float MFMA sums (in double every 4 K steps when wide)
                        np.testing.assert_allclose(got, want[i], rtol=2e-6 if variant & 2 else 1e-5)
                        # relative error where the product is nonzero (all-zero slabs leave
                        # exact zeros, which the kernels reproduce)
                        rel = np.abs(got - want[i]) / np.where(want[i] != 0, np.abs(want[i]), 1.0)
                        err[variant] = max(err.get(variant, 0.0), float(np.max(rel)))
        if esize == 4:
            # the bf16 split products keep float accuracy: no worse than twice the float
            # MFMA kernel's largest error with the same accumulation
            _say("max rel error by variant", {v: "%.2e" % e for v, e in sorted(err.items())})
            for v, e in err.items():
                base = 3 if v & 2 else 1
                if v & 16 and base in err:
                    assert e <= 2 * err[base] + 1e-7, (v, e, err[base])


@pytest.mark.parametrize("ra,rb,s", [(14, 8, 1), (12, 7, 2), (13, 12, 3), (16, 10, 0), (12, 9, 4), (11, 6, 1)])
def test_tensordot_row_form(hip, hip32, ra, rb, s):
    # joins whose outputs are long contiguous rows (the larger operand's free bits high,
    # 2^fy >= 32 contiguous y) and few shared bits: the row form of the streamed join
    rng = np.random.default_rng(ra * 31 + rb)
    _say("case row form", ra, rb, s)
    a = rng.uniform(0.5, 2.0, (2,) * ra)
    b = rng.uniform(0.5, 2.0, (2,) * rb)
    la = [int(x) for x in rng.permutation(ra)[:s]]
    lb = [int(x) for x in rng.permutation(rb)[:s]]
    fx, fy = ra - s, rb - s
    assert fx >= 6 and 5 <= fy <= 10 and s <= 4 and not (fy >= 6 and s >= 4)  # the row form's range
    want = np.tensordot(a, b, (la, lb))
    got = hip.tensordot(a, b, (la, lb)).numpy()
    np.testing.assert_allclose(got, want, rtol=1e-12)
    got32 = np.asarray(hip32.tensordot(a.astype(np.float32), b.astype(np.float32), (la, lb)).numpy(), np.float64)
    np.testing.assert_allclose(got32, np.tensordot(a.astype(np.float32).astype(np.float64),
                                                   b.astype(np.float32).astype(np.float64), (la, lb)), rtol=1e-6)


def test_cnf70_4096_slices(hip, hip32):
    # the 4096-slice random 3-CNF (70 variables, factor-Flow, max rank 25): the FP64
    # executor on the first slices against the oracle's prefix sum, then all 4096 slices
    # in both entry types
    doc = json.load(open(os.path.join(HERE, "golden", "cnf70_factor_s12.json")))
    pre = doc["meta"]["expected"]["prefix"]
    plan = plan_io.plan_from_dict(doc)
    assert len(plan.groups_to_slice) == 12
    _say("case cnf70 prefix", pre["slices"])
    got = hip.contract_sliced(plan, num_slice_limit=pre["slices"])
    assert hip.last_result.slices == pre["slices"]
    assert got == pytest.approx(pre["float64"], rel=F64_TOL)
    _say("case cnf70 full fp64")
    full64 = hip.contract_sliced(plan)
    assert hip.last_result.slices == 4096
    _say("case cnf70 full fp32s")
    full32 = hip32.contract_sliced(plan)
    assert hip32.last_result.slices == 4096
    _say("cnf70: fp64 %.17g fp32s %.17g rel %.3g" % (full64, full32, abs(full32 - full64) / abs(full64)))
    assert full32 == pytest.approx(full64, rel=F32_TOL)
    assert full64 > pre["float64"]


def test_unused_variables(hip, hip32):
    # DIMACS inputs declaring variables that no clause uses, under both planners: line
    # plans hold a rank-0 variable tensor at two leaves (the reference joins it at its first
    # bag and again at the top), and the count carries its weights twice, as the oracle's
    from test_emulator import unused_variable_plans

    plans = unused_variable_plans()
    wants = [numpy_oracle.contract_sliced(plan_io.plan_to_dict(p)) for _, p in plans]
    for api, tol in ((hip, F64_TOL), (hip32, F32_TOL)):
        for (name, plan), want in zip(plans, wants):
            _say("case unused variables", name, api.get_entry_size())
            assert api.contract_sliced(plan) == pytest.approx(want, rel=tol), name
        exe = api.prepare_batch([p for _, p in plans])
        try:
            res, _ = exe.run()
        finally:
            exe.close()
        assert [r.value for r in res] == pytest.approx(wants, rel=tol)


@pytest.mark.parametrize("early", [4, 6, 9])
def test_early_contracted_plans(hip, hip32, early):
    # plans after the reference's early contraction (dense leaves, rerooted and regrouped
    # tree, renumbered slice groups) through both executors
    from test_emulator import early_contracted_plan

    _say("case early contraction", early)
    plan, want = early_contracted_plan(early, nvars=34, nclauses=76, seed=7, slices=4)
    assert hip.contract_sliced(plan) == pytest.approx(want, rel=F64_TOL)
    assert hip32.contract_sliced(plan) == pytest.approx(want, rel=F32_TOL)


# (rank a, rank b, shared): streamed joins with fx >= 8 and fx < 8 and every y chunk,
# outer and scalar products, and the split-K form (s > 12, and s >= 8 with few outputs)
C_TENSORDOT_SHAPES = [(0, 0, 0), (3, 0, 0), (0, 5, 0), (4, 4, 4), (6, 3, 0), (9, 5, 3), (12, 10, 6), (14, 14, 10),
                      (13, 9, 1), (11, 12, 7), (20, 20, 20), (16, 16, 14), (18, 10, 9), (15, 14, 13), (10, 16, 3)]


@pytest.mark.parametrize("et", ["f64", "f32s"])
def test_c_tensordot_matches_numpy(et):
    # the C entry point to_tensordot (numpy.tensordot on the device, workspace from the
    # caller) against numpy in float64
    import torch

    from ctensordot import c_tensordot

    from tensororder_amd import _lib

    lib = _lib.load()
    stream = torch.cuda.Stream()
    entry_type = _lib.TO_F64 if et == "f64" else _lib.TO_F32S
    rng = np.random.default_rng(17)
    for ra, rb, s in C_TENSORDOT_SHAPES:
        _say("case c tensordot", et, ra, rb, s)
        a = rng.uniform(0.5, 2.0, (2,) * ra)
        b = rng.uniform(0.5, 2.0, (2,) * rb)
        la = [int(x) for x in rng.permutation(ra)[:s]]
        lb = [int(x) for x in rng.permutation(rb)[:s]]
        if et == "f32s":  # the float32 inputs the device sees
            a = a.astype(np.float32).astype(np.float64)
            b = b.astype(np.float32).astype(np.float64)
        want = np.tensordot(a, b, (la, lb))
        got = c_tensordot(lib, stream, a, b, (la, lb), entry_type)
        np.testing.assert_allclose(got, want, rtol=1e-12 if et == "f64" else 1e-6)


def test_batch_admission_waves(hip32, hip):
    # the 512-instance fixture under a device budget of about a third of its arena: the
    # instances are admitted in waves (one graph each, one shared arena), and every count
    # matches the oracle's
    from tensororder_amd import batch

    doc = json.load(open(os.path.join(HERE, "golden", "vc_batch512.json")))
    plans, wants = plan_io.batch_plans(doc)
    for api, tol in ((hip32, F32_TOL), (hip, F64_TOL)):
        instances = []
        for plan in plans:
            _total, tree, tables, _groups = api._plan_inputs(plan, None)
            instances.append((tree, tables))
        per = batch.analyse_batch(instances, api._entry_type)
        need = [p["region"] + p["fixed"] for p in per]
        limit = max(max(need), sum(need) // 3)
        _say("case batch admission waves", api.get_entry_size(), "limit", limit, "of", sum(need))
        api.add_argument("mem_limit", limit)
        try:
            exe = api.prepare_batch(plans)
        finally:
            api.add_argument("mem_limit", None)
        try:
            assert len(exe.progs) >= 2
            assert sum(p.count for p in exe.progs) == len(plans)
            for rep in range(2):  # graph replays
                res, lat = exe.run()
                assert [r.value for r in res] == pytest.approx(wants, rel=tol)
                assert all(l >= 0 for l in lat)
        finally:
            exe.close()


@pytest.mark.parametrize("et", ["f64", "f32s"])
def test_device_leaf_tables_match_host_build(hip, hip32, et):
    # every OR / variable leaf kind (plain and output-indexed OR, factored OR and variable
    # pieces, rank 0..5, weights unweighted / cachet / mcc) under several slicings: the
    # tables the leaf build kernel writes equal, byte for byte, the host's network build
    # reordered as tensor.py:95-119 reorders it (float32 encoding in the scaled executor)
    import ctypes

    import leafcases
    import torch

    from tensororder_amd import _lib, schedule
    from tensororder_amd.hip_api import leaf_spec_table

    api = hip if et == "f64" else hip32
    scaled = et == "f32s"
    dt = np.float32 if scaled else np.float64
    cases = leafcases.cases()
    _say("case device leaves", et, len(cases))
    offs, off, descs, wants, block0 = [], 0, [], [], 0
    for name, t, lookup, ngroups, rows, coef in cases:
        tab = leaf_spec_table(t, lookup, ngroups)
        v0, v1, F, e = tab.spec.stored(scaled)
        want = schedule.leaf_encoding(rows)[0].astype(np.float32) if scaled else rows
        wants.append((name, want))
        offs.append(off)
        descs.append((tab.spec, v0, v1))
        off += (tab.size + 63) // 64 * 64
    esize = 4 if scaled else 8
    buf = torch.full((off,), float("nan"), dtype=torch.float32 if scaled else torch.float64, device="cuda")
    structs = []
    for (spec, v0, v1), o in zip(descs, offs):
        bd = spec.desc(buf.data_ptr() + o * esize, v0, v1)
        bd.block0 = block0
        block0 += bd.nblocks
        structs.append(bd)
    blob = torch.frombuffer(bytearray(b"".join(bytes(s) for s in structs)), dtype=torch.uint8).cuda()
    torch.cuda.synchronize()
    L = _lib.Launch(op=_lib.TO_OP_LEAF_BUILD, count=len(structs), grid=block0, descs=blob.data_ptr())
    arr = _lib.launch_array([L])
    _lib.check(api._lib.to_run(ctypes.addressof(arr), 1, api._entry_type, ctypes.c_void_p(api._stream.cuda_stream)),
               "to_run")
    api._stream.synchronize()
    host = buf.cpu().numpy()
    for (name, want), o in zip(wants, offs):
        got = host[o:o + want.size].reshape(want.shape)
        assert got.tobytes() == np.ascontiguousarray(want, dtype=dt).tobytes(), name


def test_device_and_host_leaves_agree(hip, hip32):
    # the same sliced plans with leaves written on the device and built on the host give
    # the same counts, bit for bit
    for name in ("cnf20_factor_s4", "vc22_weighted_line_s3"):
        doc = json.load(open(os.path.join(HERE, "golden", name + ".json")))
        for api in (hip, hip32):
            _say("case device vs host leaves", name, api.get_entry_size())
            plan = plan_io.plan_from_dict(doc)
            api.add_argument("device_leaves", False)
            try:
                host = api.contract_sliced(plan)
            finally:
                api.add_argument("device_leaves", True)
            dev = api.contract_sliced(plan)
            assert api.last_stats["device_built_leaves"] == len(doc["tensors"])
            assert dev == host
