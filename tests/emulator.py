"""Test-only numpy emulator of the launch records in ``include/tensororder_hip.h``.

It executes a compiled Program's records against a simulated device heap so the schedule
compiler (bit placement, memory plan, batching, leaf gather, root reduction) can be
checked on a machine without a GPU.  It restates the kernels' documented semantics, not
their code, and is never used by the product.
"""

import ctypes
import math

import numpy as np

from tensororder_amd import _lib

HEAP_BASE = 1 << 40


class Heap:
    """A flat byte heap handing out fake device addresses."""

    def __init__(self, nbytes=1 << 28):
        self.mem = np.zeros(nbytes, dtype=np.uint8)
        self.top = 0

    def alloc(self, nbytes):
        start = (self.top + 255) // 256 * 256
        self.top = start + int(nbytes)
        if self.top > self.mem.size:
            raise MemoryError("emulated heap exhausted")
        heap = self

        class Buf:
            def data_ptr(self_inner):
                return HEAP_BASE + start

            def __getitem__(self_inner, sl):
                return heap.mem[start:][sl]

        return Buf()

    def view(self, addr, dtype, count):
        off = addr - HEAP_BASE
        nbytes = np.dtype(dtype).itemsize * count
        return self.mem[off:off + nbytes].view(dtype)

    def write(self, buf, data):
        off = buf.data_ptr() - HEAP_BASE
        self.mem[off:off + len(data)] = np.frombuffer(data, dtype=np.uint8)

    def read_struct(self, cls, addr):
        off = addr - HEAP_BASE
        return cls.from_buffer_copy(self.mem[off:off + ctypes.sizeof(cls)].tobytes())


def pdep_array(vals, pos, n):
    out = np.zeros(len(vals), dtype=np.int64)
    for i in range(n):
        out |= ((vals >> i) & 1) << int(pos[i])
    return out


def _scale_exp(bits):
    m = float(np.array([bits], dtype=np.uint32).view(np.float32)[0])
    return math.frexp(m)[1] if m > 0 else 0


def _sum_mexp(terms):
    """Sum of values m * 2^e given as (m, e) pairs, as a normalised (mantissa, exponent)."""
    terms = [(m, e) for m, e in terms if m != 0.0]
    if not terms:
        return 0.0, 0
    top = max(e for _, e in terms)
    total = sum(math.ldexp(m, e - top) for m, e in terms)
    m, e = math.frexp(total)
    return m, (e + top if m else 0)


class Emulator:
    def __init__(self, heap, entry_type):
        self.clock = 0
        self.heap = heap
        self.scaled = entry_type == _lib.TO_F32S
        self.dt = np.float32 if self.scaled else np.float64
        self.maps_checked = 0
        self.wave_joins = 0
        self.gemm_partials = set()
        self.split_gemms = 0

    def recs(self, addr, count):
        raw = self.heap.view(addr, np.uint8, 16 * count).reshape(count, 16)
        return raw

    def get_rec(self, addr, idx):
        r = self.heap.read_struct(_lib.ScaleRec, addr + 16 * idx)
        return r.exp, r.maxbits

    def set_rec(self, addr, idx, exp=None, maxbits=None):
        off = addr - HEAP_BASE + 16 * idx
        if exp is not None:
            self.heap.mem[off:off + 8] = np.frombuffer(np.int64(exp).tobytes(), np.uint8)
        if maxbits is not None:
            cur = int(self.heap.mem[off + 8:off + 12].view(np.uint32)[0])
            new = max(cur, int(maxbits))
            self.heap.mem[off + 8:off + 12] = np.frombuffer(np.uint32(new).tobytes(), np.uint8)

    TABLE_DESCS = {1: "LeafDesc", 2: "JoinDesc", 3: "JoinDesc", 4: "JoinDesc", 9: "PermDesc", 10: "GemmDesc",
                   13: "LeafBuildDesc"}

    def check_block_map(self, L):
        """A launch's block -> descriptor map (aux1) must name, for every block, the last
        descriptor whose block0 is at or below it -- what the kernels' binary search finds."""
        cls = getattr(_lib, self.TABLE_DESCS[L.op])
        size = ctypes.sizeof(cls)
        b0 = np.array([self.heap.read_struct(cls, L.descs + i * size).block0 for i in range(L.count)])
        assert b0[0] == 0 and np.all(np.diff(b0) >= 0)
        bmap = self.heap.view(L.aux1, np.int32, L.grid)
        want = np.searchsorted(b0, np.arange(L.grid), side="right") - 1
        assert np.array_equal(bmap, want), "block map disagrees with block0"
        self.maps_checked += 1

    def run(self, launches):
        for L in launches:
            if L.op in self.TABLE_DESCS and L.aux1:
                self.check_block_map(L)
            getattr(self, "op_%d" % L.op)(L)

    # leaf gather
    def op_1(self, L):
        for i in range(L.count):
            d = self.heap.read_struct(_lib.LeafDesc, L.descs + i * ctypes.sizeof(_lib.LeafDesc))
            begin = int(self.heap.view(L.aux3, np.int64, 1)[0]) if L.aux3 else 0
            E = 1 << d.rank
            coef = self.heap.view(d.coef, np.int64, max(1, d.ngroups))
            for b in range(d.nbatch):
                sidx = begin + b
                row = sum(int(coef[g]) * ((sidx >> (d.ngroups - 1 - g)) & 1) for g in range(d.ngroups))
                src = self.heap.view(d.src + row * E * self.dt().itemsize, self.dt, E)
                dst = self.heap.view(d.dst + b * E * self.dt().itemsize, self.dt, E)
                dst[:] = src
                if self.scaled and d.rec:
                    one = int(np.array([1.0], np.float32).view(np.uint32)[0])
                    self.set_rec(d.rec, b, exp=d.exp0)
                    off = d.rec - HEAP_BASE + 16 * b + 8
                    self.heap.mem[off:off + 4] = np.frombuffer(np.uint32(one).tobytes(), np.uint8)

    def _join(self, d, write_partial=False, finish=False):
        dt = self.dt
        fx, fy, s = d.fx, d.fy, d.s
        f = np.arange(1 << fx, dtype=np.int64)
        y = np.arange(1 << fy, dtype=np.int64)
        k = np.arange(1 << s, dtype=np.int64)
        xf, xs = pdep_array(f, d.xf_pos, fx), pdep_array(k, d.xs_pos, s)
        yf, ys = pdep_array(y, d.yf_pos, fy), pdep_array(k, d.ys_pos, s)
        cidx = (f[:, None] << d.shf) | (y[None, :] << d.shy)
        esz = np.dtype(dt).itemsize
        for b in range(d.nbatch):
            X = self.heap.view(d.x + b * d.xb * esz, dt, int((xf.max(initial=0) + xs.max(initial=0)) + 1))
            Y = self.heap.view(d.y + b * d.yb * esz, dt, int((yf.max(initial=0) + ys.max(initial=0)) + 1))
            sx = sy = 1.0
            ex = ey = 0
            if self.scaled:
                e0, m0 = self.get_rec(d.xs, b * d.xsb)
                e1, m1 = self.get_rec(d.ys, b * d.ysb)
                sx, ex = 2.0 ** -_scale_exp(m0), e0 + _scale_exp(m0)
                sy, ey = 2.0 ** -_scale_exp(m1), e1 + _scale_exp(m1)
            Xm = X[xf[:, None] + xs[None, :]].astype(np.float64) * sx
            Ym = Y[yf[:, None] + ys[None, :]].astype(np.float64) * sy
            Cm = (Xm @ Ym.T).astype(dt)
            C = self.heap.view(d.c + b * d.cb * esz, dt, int(cidx.max() + 1))
            C[cidx] = Cm
            if self.scaled:
                m = float(np.max(np.abs(Cm))) if Cm.size else 0.0
                self.set_rec(d.cs, b, exp=ex + ey,
                             maxbits=int(np.array([m], np.float32).view(np.uint32)[0]))

    def op_2(self, L):
        valid = {y | f for y in (1, 2, 4, 8, 16) for f in (0, _lib.TO_STREAM_VEC)} | {_lib.TO_STREAM_ROWS}
        if L.variant == _lib.TO_STREAM_WAVE:
            # one wave per descriptor, four per block: the grid covers every descriptor
            assert L.grid * 4 >= L.count and L.lds == 0 and not L.aux1
            self.wave_joins += L.count
        for i in range(L.count):
            d = self.heap.read_struct(_lib.JoinDesc, L.descs + i * ctypes.sizeof(_lib.JoinDesc))
            if L.variant == _lib.TO_STREAM_WAVE:
                assert d.kvar == _lib.TO_STREAM_WAVE
            else:
                # a launch of one form carries it in every descriptor; a mixed launch names each
                assert d.kvar in valid and (L.variant == _lib.TO_STREAM_ANY or d.kvar == L.variant)
            self._join(d)

    def op_3(self, L):  # split-K partials: the emulator computes the join at FINISH
        pass

    def op_4(self, L):
        for i in range(L.count):
            d = self.heap.read_struct(_lib.JoinDesc, L.descs + i * ctypes.sizeof(_lib.JoinDesc))
            if d.partial in self.gemm_partials:
                self._finish_from_partials(d)
            else:
                self._join(d)

    def _finish_from_partials(self, d):
        """TO_OP_JOIN_FINISH over the partial sums a split-K product wrote: output o =
        (batch << (fx + fy)) | (f << fy) | y adds partial[o * nsplit + sp] over sp."""
        fx, fy, dt = d.fx, d.fy, self.dt
        nout = d.nbatch << (fx + fy)
        P = self.heap.view(d.partial, np.float64, nout * d.nsplit).reshape(nout, d.nsplit)
        sums = P.sum(axis=1).astype(dt).reshape(d.nbatch, 1 << fx, 1 << fy)
        f = np.arange(1 << fx, dtype=np.int64)
        y = np.arange(1 << fy, dtype=np.int64)
        cidx = (f[:, None] << d.shf) | (y[None, :] << d.shy)
        esz = np.dtype(dt).itemsize
        for b in range(d.nbatch):
            C = self.heap.view(d.c + b * d.cb * esz, dt, int(cidx.max() + 1))
            C[cidx] = sums[b]
            if self.scaled:
                e0, m0 = self.get_rec(d.xs, b * d.xsb)
                e1, m1 = self.get_rec(d.ys, b * d.ysb)
                m = float(np.max(np.abs(sums[b])))
                self.set_rec(d.cs, b, exp=e0 + _scale_exp(m0) + e1 + _scale_exp(m1),
                             maxbits=int(np.array([m], np.float32).view(np.uint32)[0]))

    def op_5(self, L):
        raw = self.recs(L.aux0, L.iaux0)
        raw[:] = 0

    def op_6(self, L):
        params = self.heap.view(L.aux3, np.int64, 2)
        begin, end = int(params[0]), int(params[1])
        nvalid = max(0, min(int(L.iaux0), end - begin))
        vals = self.heap.view(L.aux0, self.dt, max(1, int(L.iaux0)))
        acc = self.heap.read_struct(_lib.Accum, L.aux2)
        terms = [(acc.mant, acc.exp)]
        for b in range(nvalid):
            m, e = math.frexp(float(vals[b]))
            if self.scaled:
                e += self.get_rec(L.aux1, b)[0]
            terms.append((m, e))
        m, e = _sum_mexp(terms)
        acc.mant, acc.exp, acc.count = m, e, acc.count + nvalid
        off = L.aux2 - HEAP_BASE
        self.heap.mem[off:off + ctypes.sizeof(_lib.Accum)] = np.frombuffer(bytes(acc), np.uint8)

    def op_9(self, L):
        for i in range(L.count):
            d = self.heap.read_struct(_lib.PermDesc, L.descs + i * ctypes.sizeof(_lib.PermDesc))
            n = 1 << d.rank
            out_pos = {}
            for j in range(d.nu):
                out_pos[d.u_in[j]] = d.v_out[[k for k in range(d.nu) if d.v2u[k] == j][0]]
            for j in range(d.nrest):
                out_pos[d.rest_in[j]] = d.rest_out[j]
            idx = np.arange(n, dtype=np.int64)
            dst_idx = np.zeros(n, dtype=np.int64)
            for p, q in out_pos.items():
                dst_idx |= ((idx >> p) & 1) << q
            esz = self.dt().itemsize
            for b in range(d.nbatch):
                src = self.heap.view(d.in_ + b * d.inb * esz, self.dt, n)
                dst = self.heap.view(d.out + b * d.outb * esz, self.dt, n)
                dst[dst_idx] = src

    def op_10(self, L):
        for i in range(L.count):
            d = self.heap.read_struct(_lib.GemmDesc, L.descs + i * ctypes.sizeof(_lib.GemmDesc))
            M, N, K = 1 << d.mlog, 1 << d.nlog, 1 << d.klog
            esz = self.dt().itemsize
            m = np.arange(M, dtype=np.int64)
            nn = np.arange(N, dtype=np.int64)
            cidx = m[:, None] * d.ldcm + nn[None, :] * d.ldcn
            if d.lsplit:
                # split K: the parts' partial sums, double[o][2^lsplit] with
                # o = (b << (mlog + nlog)) | (m << nlog) | n
                S = 1 << d.lsplit
                P = self.heap.view(d.c, np.float64, (d.nbatch << (d.mlog + d.nlog)) * S)
                P = P.reshape(d.nbatch, M, N, S)
                for b in range(d.nbatch):
                    A = self.heap.view(d.a + b * d.ab * esz, self.dt, M * K).reshape(M, K).astype(np.float64)
                    B = self.heap.view(d.b + b * d.bb * esz, self.dt, N * K).reshape(N, K).astype(np.float64)
                    if self.scaled:
                        A = A * 2.0 ** -_scale_exp(self.get_rec(d.as_, b * d.asb)[1])
                        B = B * 2.0 ** -_scale_exp(self.get_rec(d.bs, b * d.bsb)[1])
                    for sp in range(S):
                        ks = slice(sp * (K // S), (sp + 1) * (K // S))
                        P[b, :, :, sp] = A[:, ks] @ B[:, ks].T
                self.gemm_partials.add(d.c)
                self.split_gemms += 1
                continue
            for b in range(d.nbatch):
                A = self.heap.view(d.a + b * d.ab * esz, self.dt, M * K).reshape(M, K).astype(np.float64)
                B = self.heap.view(d.b + b * d.bb * esz, self.dt, N * K).reshape(N, K).astype(np.float64)
                ea = eb = 0
                if self.scaled:
                    e0, m0 = self.get_rec(d.as_, b * d.asb)
                    e1, m1 = self.get_rec(d.bs, b * d.bsb)
                    A = A * 2.0 ** -_scale_exp(m0)
                    B = B * 2.0 ** -_scale_exp(m1)
                    ea, eb = e0 + _scale_exp(m0), e1 + _scale_exp(m1)
                Cm = (A @ B.T).astype(self.dt)
                C = self.heap.view(d.c + b * d.cb * esz, self.dt, int(cidx.max() + 1))
                C[cidx] = Cm
                if self.scaled:
                    mx = float(np.max(np.abs(Cm)))
                    self.set_rec(d.cs, b, exp=ea + eb, maxbits=int(np.array([mx], np.float32).view(np.uint32)[0]))

    def op_11(self, L):
        for i in range(L.count):
            d = self.heap.read_struct(_lib.RootDesc, L.descs + i * ctypes.sizeof(_lib.RootDesc))
            m, e = math.frexp(float(self.heap.view(d.root, self.dt, 1)[0]))
            if self.scaled:
                e += self.get_rec(d.rec, 0)[0]
            acc = _lib.Accum(mant=m, exp=e if m else 0, count=1, pad=self.clock)
            self.clock += 1
            off = d.acc - HEAP_BASE
            self.heap.mem[off:off + ctypes.sizeof(_lib.Accum)] = np.frombuffer(bytes(acc), np.uint8)

    def op_12(self, L):
        raw = self.heap.view(L.aux2, np.uint8, L.iaux0 * ctypes.sizeof(_lib.Accum))
        raw[:] = 0
        self.clock = 1

    def op_13(self, L):  # leaf build: the documented table of each OR / variable leaf
        size = ctypes.sizeof(_lib.LeafBuildDesc)
        for i in range(L.count):
            d = self.heap.read_struct(_lib.LeafBuildDesc, L.descs + i * size)
            r = d.rank
            e = np.arange(1 << r, dtype=np.int64)
            full = pdep_array(e, list(d.pos), r)
            if d.kind == _lib.TO_LEAF_VAR:
                val = np.where((full == (1 << r) - 1) & (r > 0), d.val1,
                               np.where((full == 0) | (r == 0), d.val0, 0.0))
            else:
                lit = full ^ d.negmask
                if d.out_bit < 0:
                    one = lit != 0
                else:
                    ob = 1 << d.out_bit
                    one = ((lit & ob) != 0) == ((lit & ~ob) != 0)
                val = np.where(one, d.val1, d.val0)
            self.heap.view(d.dst, self.dt, 1 << r)[:] = val

    def op_7(self, L):
        params = self.heap.view(L.aux3, np.int64, 2)
        params[0] += L.iaux0


def emulate_plan(tree, tables, num_groups, total, entry_type, W, regroup=True):
    """Compile and emulate; returns (value, program)."""
    from tensororder_amd import schedule

    heap = Heap()
    prog = schedule.compile_program(tree, tables, num_groups, entry_type, W, heap.alloc, regroup=regroup)
    heap.write(prog.table, prog.table_host.tobytes())
    heap.write(prog.coef, prog.coef_host.tobytes())
    heap.write(prog.blob, prog.blob_host.tobytes())
    if entry_type == _lib.TO_F32S:
        heap.write(prog.recs, schedule.rec_init_bytes(prog))
    heap.write(prog.params, np.array([0, total, 0, 1], dtype=np.int64).tobytes())
    heap.write(prog.accum, bytes(_lib.Accum()))
    emu = Emulator(heap, entry_type)
    emu.run(prog.build)
    emu.run(prog.prologue)
    if not prog.root_const:
        for _ in range(-(-total // W)):
            emu.run(prog.wave)
    acc = heap.read_struct(_lib.Accum, prog.accum.data_ptr())
    fm, fe = getattr(prog, "leaf_factor", (1.0, 0))
    return math.ldexp(acc.mant * fm, acc.exp + fe), prog


def plan_tables(plan, api_like, device_leaves=True):
    """(tree, tables, num_groups, total) exactly as HipAPI.contract_sliced builds them (OR
    and variable leaves described for the device to write, with ``device_leaves``)."""
    from tensororder_amd import schedule
    from tensororder_amd.hip_api import leaf_spec_table

    groups = [g for g in plan.groups_to_slice if len(g) > 0]
    total, sequences, _ = plan.network.get_tensor_slices(plan.groups_to_slice, api_like.create_tensor)
    tree = plan.network.remove_sliced_indices_from(plan.tree, plan.groups_to_slice)
    _total, lookups = plan.network.slice_lookups(plan.groups_to_slice)
    assert _total == total
    tables = {}
    for t, seq in enumerate(sequences):
        spec_table = leaf_spec_table(plan.network[t], lookups[t], len(groups)) if device_leaves else None
        if spec_table is not None:
            tables[t] = spec_table
            continue
        reordered, lookup, shape = seq.reordered_tensor(api_like)
        reordered = np.asarray(reordered, dtype=np.float64)
        tables[t] = schedule.LeafTable(reordered.reshape(reordered.shape[0], -1), lookup, len(shape))
    return tree, tables, len(groups), total


class HostOps:
    """The host-side leaf helpers of HipAPI (numpy), without needing a GPU."""

    @staticmethod
    def create_tensor(shape, default_value=None):
        if default_value is None:
            return np.empty(shape, dtype=np.float64)
        return np.full(shape, default_value, dtype=np.float64)

    @staticmethod
    def reshape(t, shape):
        return np.reshape(t, shape)

    @staticmethod
    def transpose(t, perm):
        return np.transpose(t, axes=perm)


def emulate_batch(instances, entry_type, budget=None):
    """Compile a batch of (tree, tables) and emulate it; returns the per-instance values and
    the program (with ``budget``: admission waves sharing one arena, run in order, and the
    list of wave programs)."""
    from tensororder_amd import batch

    heap = Heap(1 << 29)
    if budget is None:
        progs = [batch.compile_batch(instances, entry_type, heap.alloc)]
    else:
        per = batch.analyse_batch(instances, entry_type)
        waves = batch.admission_waves(per, budget)
        assert waves is not None
        arena = heap.alloc(max(batch.wave_arena_bytes(per, w) for w in waves))
        progs = [batch.compile_batch(None, entry_type, heap.alloc, per=[per[i] for i in w], arena=arena)
                 for w in waves]
    out = []
    for prog in progs:
        heap.write(prog.table, prog.table_host.tobytes())
        heap.write(prog.blob, prog.blob_host.tobytes())
        if entry_type == _lib.TO_F32S:
            heap.write(prog.recs, batch.rec_init_bytes(prog))
    emu = Emulator(heap, entry_type)
    for prog in progs:
        emu.run(prog.build)
    for prog in progs:
        emu.run(prog.launches)
        for i in range(prog.count):
            acc = heap.read_struct(_lib.Accum, prog.accum.data_ptr() + i * ctypes.sizeof(_lib.Accum))
            fm, fe = prog.leaf_factors[i]
            out.append(math.ldexp(acc.mant * fm, acc.exp + fe))
    progs[0].maps_checked = emu.maps_checked
    progs[0].wave_joins = emu.wave_joins
    progs[0].split_gemms = emu.split_gemms
    return out, (progs[0] if budget is None else progs)
