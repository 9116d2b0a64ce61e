"""The MI355X ``tensor_library``: a drop-in ``BaseTensorAPI`` backed by HIP kernels.

Mirrors the plugin interface the reference's ``execution.run`` drives
(``tensor_network/tensor_apis/base_api.py:8-28``; registry ``tensor_network/__init__.py:12-16``):

* ``add_argument(key, value)``: ``entry_type`` (``float64`` = plain FP64 executor,
  ``float32`` = FP32 with per-tensor power-of-two scaling), ``thread_limit`` (accepted;
  host threads do no arithmetic here), ``max_batch``, ``mem_limit`` (bytes of device
  memory the executor may use), ``graph`` (capture waves in a HIP graph), ``regroup``,
  ``autotune`` (time the launch shapes of new product/permute shapes before compiling);
  any other key raises ``ValueError`` (``base_api.py:9-12``).
* ``contract_sliced(plan, num_slice_limit)``: all slices of the plan, batched in waves on
  the device, summed in FP64 (mantissa, exponent) -- the override pattern of
  ``jax_apis.py:316-322``; returns a Python float.
* ``contract(network, tree)``: one (possibly sliced) network; returns a 0-d array.
* ``create_tensor`` / ``reshape`` / ``transpose`` build leaf values on the host (leaves
  are tiny), ``tensordot`` contracts two tensors on the device (the per-join primitive
  ``network.identify`` calls, ``tensor_network.pyx:152-154``).

Device memory is allocated through PyTorch; the arithmetic runs only in the HIP library.
If the library or the GPU is missing, construction raises.
"""

import ctypes
import math

import numpy as np

from . import _lib, schedule
from .api import BaseTensorAPI, OutOfMemoryError
from .tree import TreeContext


class DeviceTensor:
    """A dense tensor in device memory (entry type of its API), optionally with a
    power-of-two scale record (FP32 mode)."""

    def __init__(self, data, shape, rec=None, api=None):
        self.data = data
        self.shape = tuple(shape)
        self.rec = rec
        self._api = api

    @property
    def rank(self):
        return len(self.shape)

    def numpy(self):
        vals = self.data.double().cpu().numpy().reshape(self.shape)
        if self.rec is not None:
            exp = int(self.rec.cpu().numpy().view(np.int64)[0])
            vals = np.ldexp(vals, exp)
        return vals

    def __getitem__(self, item):
        return self.numpy()[item]


class Result:
    """A count as FP64 mantissa and int64 exponent (value = mant * 2^exp)."""

    def __init__(self, mant, exp, slices):
        self.mant, self.exp, self.slices = float(mant), int(exp), int(slices)

    @property
    def value(self):
        try:
            return math.ldexp(self.mant, self.exp)
        except OverflowError:
            return math.copysign(math.inf, self.mant)

    @property
    def log10(self):
        if self.mant == 0:
            return -math.inf
        return math.log10(abs(self.mant)) + self.exp * math.log10(2.0)


class Executable:
    """A compiled plan resident on the device.  ``run()`` is one full contraction: reset
    the accumulator, prologue (slice-invariant subtrees), then every wave of slices,
    replayed from HIP graphs when the API's ``graph`` option is on."""

    def __init__(self, api, prog, total_slices):
        self.api = api
        self.prog = prog
        self.total = total_slices
        self.nwaves = 0 if prog.root_const else -(-total_slices // prog.nbatch)
        begin = _lib.Launch(op=_lib.TO_OP_STEP_BEGIN, grid=1, aux2=prog.accum.data_ptr(),
                            aux3=prog.params.data_ptr())
        self.step_launches = [begin] + list(prog.prologue)
        self.wave_launches = list(prog.wave)
        self._graphs = []
        self._pro_graph = None
        self._wave_graph = None
        if api._use_graph:
            self._pro_graph = self._capture(self.step_launches)
            if self.nwaves > 1:
                self._wave_graph = self._capture(self.wave_launches)

    def _capture(self, launches):
        arr = _lib.launch_array(launches)
        graph = ctypes.c_void_p()
        _lib.check(self.api._lib.to_graph_create(ctypes.addressof(arr), len(launches), self.api._entry_type,
                                                 ctypes.c_void_p(self.api._stream.cuda_stream),
                                                 ctypes.byref(graph)), "to_graph_create")
        self._graphs.append(graph)
        return graph

    def _launch_graph(self, graph):
        _lib.check(self.api._lib.to_graph_launch(graph, ctypes.c_void_p(self.api._stream.cuda_stream)),
                   "to_graph_launch")

    def enqueue(self):
        """Queue one full contraction on the API's stream (no synchronisation)."""
        if self._pro_graph is not None:
            self._launch_graph(self._pro_graph)
        else:
            self.api._run(self.step_launches)
        for _ in range(self.nwaves):
            if self._wave_graph is not None:
                self._launch_graph(self._wave_graph)
            else:
                self.api._run(self.wave_launches)

    def result(self):
        """Synchronise and read the (mantissa, exponent) slice sum."""
        api = self.api
        api.synchronize()
        with api._torch.cuda.stream(api._stream):
            acc_bytes = self.prog.accum[: ctypes.sizeof(_lib.Accum)].cpu().numpy().tobytes()
        acc = _lib.Accum.from_buffer_copy(acc_bytes)
        fm, fe = getattr(self.prog, "leaf_factor", (1.0, 0))
        res = Result(acc.mant * fm, acc.exp + fe, acc.count)
        api.last_result = res
        api.last_stats = dict(self.prog.stats, nbatch=self.prog.nbatch, waves=self.nwaves, slices=self.total)
        return res

    def run(self):
        self.enqueue()
        return self.result()

    def close(self):
        for g in self._graphs:
            self.api._lib.to_graph_destroy(g)
        self._graphs = []
        self._pro_graph = self._wave_graph = None


class BatchExecutable:
    """A batch of unsliced plans resident on the device, admitted in waves sized to the
    device budget: one grouped schedule (one HIP graph) per wave, the waves run in order
    through one shared arena.  ``run()`` contracts all of them and returns their Results and
    per-instance completion times (s from the batch start, device clock)."""

    def __init__(self, api, progs, arena=None):
        self.api = api
        self.progs = list(progs)
        self.prog = self.progs[0]
        self.arena = arena
        self.count = sum(p.count for p in self.progs)
        self._graphs = []
        if api._use_graph:
            for prog in self.progs:
                arr = _lib.launch_array(prog.launches)
                graph = ctypes.c_void_p()
                _lib.check(api._lib.to_graph_create(ctypes.addressof(arr), len(prog.launches), api._entry_type,
                                                    ctypes.c_void_p(api._stream.cuda_stream), ctypes.byref(graph)),
                           "to_graph_create")
                self._graphs.append(graph)
        khz = ctypes.c_int(0)
        _lib.check(api._lib.to_wall_clock_khz(ctypes.byref(khz)), "to_wall_clock_khz")
        self.clock_hz = khz.value * 1e3

    @property
    def launches(self):
        return [L for prog in self.progs for L in prog.launches]

    @property
    def work(self):
        return [w for prog in self.progs for w in prog.work]

    def enqueue(self):
        if self._graphs:
            for graph in self._graphs:
                _lib.check(self.api._lib.to_graph_launch(graph, ctypes.c_void_p(self.api._stream.cuda_stream)),
                           "to_graph_launch")
        else:
            for prog in self.progs:
                self.api._run(prog.launches)

    def results(self):
        api = self.api
        api.synchronize()
        size = ctypes.sizeof(_lib.Accum)
        res, ends = [], []
        with api._torch.cuda.stream(api._stream):
            start = int(self.progs[0].start[:8].cpu().numpy().view(np.int64)[0])
            for prog in self.progs:
                raw = prog.accum[: prog.count * size].cpu().numpy().tobytes()
                accs = [_lib.Accum.from_buffer_copy(raw[i * size:(i + 1) * size]) for i in range(prog.count)]
                res += [Result(a.mant * fm, a.exp + fe, a.count) for a, (fm, fe) in zip(accs, prog.leaf_factors)]
                ends += [a.pad for a in accs]
        lat = [(e - start) / self.clock_hz for e in ends]
        return res, lat

    def run(self):
        self.enqueue()
        return self.results()

    def close(self):
        for graph in self._graphs:
            self.api._lib.to_graph_destroy(graph)
        self._graphs = []


def leaf_spec_table(tensor, lookup, ngroups):
    """The LeafTable of an OR or variable leaf whose table the device writes, with its axes
    reordered and its row lookup exactly as ``SliceSequence.reordered_tensor`` gives them
    (``tensor.py:95-119``); None for other leaves."""
    from .network import OrTensor, VariableTensor

    if isinstance(tensor, VariableTensor):
        w_neg, w_pos = tensor.weights
        kind, kw = _lib.TO_LEAF_VAR, dict(w_neg=w_neg, w_pos=w_pos)
    elif isinstance(tensor, OrTensor):
        kind, kw = _lib.TO_LEAF_OR, dict(positive=tensor.literals_positive, output_index=tensor.output_index)
    else:
        return None
    rank = tensor.rank
    sliced = [i for i in range(rank) if i in lookup]
    unsliced = [i for i in range(rank) if i not in lookup]
    coef = [0] * ngroups
    offset = 1
    for i in reversed(sliced):
        coef[lookup[i]] += offset
        offset *= 2
    spec = schedule.LeafSpec(kind, rank, sliced + unsliced, **kw)
    return schedule.LeafTable(None, coef, len(unsliced), spec)


class HipAPI(BaseTensorAPI):
    def __init__(self, device=0):
        import torch

        self._torch = torch
        self._lib = _lib.load()
        if not torch.cuda.is_available():
            raise RuntimeError("HipAPI needs a ROCm GPU; none is visible")
        self._device = torch.device("cuda", device)
        self._stream = torch.cuda.Stream(device=self._device)
        self._entry_type = _lib.TO_F64
        self._thread_limit = None
        self._max_batch = 1 << 16
        self._mem_limit = None
        self._use_graph = True
        self._regroup = True
        self._autotune = False
        self._device_leaves = True
        # launch shapes tuned on this GPU model (autotune.py), used where a shape is known
        from . import autotune

        schedule.TUNED.update(autotune.load())
        self.last_result = None
        self.last_stats = {}

    # -- plugin interface ---------------------------------------------------------

    def add_argument(self, key, value):
        if key == "entry_type":
            types = {"float64": _lib.TO_F64, "float32": _lib.TO_F32S}
            if value not in types:
                raise ValueError("Unknown hip type %s" % value)
            self._entry_type = types[value]
        elif key == "thread_limit":
            self._thread_limit = value
        elif key == "max_batch":
            self._max_batch = max(1, int(value))
        elif key == "mem_limit":
            self._mem_limit = None if value is None else int(value)
        elif key == "graph":
            self._use_graph = bool(value)
        elif key == "regroup":
            self._regroup = bool(value)
        elif key == "autotune":
            self._autotune = bool(value)
        elif key == "device_leaves":
            self._device_leaves = bool(value)
        else:
            super().add_argument(key, value)

    def get_entry_size(self):
        return 8 if self._entry_type == _lib.TO_F64 else 4

    def warm(self):
        a = self.create_tensor((2,), 2)
        self.tensordot(a, a, ([0], [0]))
        self.synchronize()

    def create_tensor(self, shape, default_value=None):
        if default_value is None:
            return np.empty(shape, dtype=np.float64)
        return np.full(shape, default_value, dtype=np.float64)

    def reshape(self, tensor, new_shape):
        return np.reshape(tensor, new_shape)

    def transpose(self, tensor, perm):
        return np.transpose(tensor, axes=perm)

    # -- device helpers -----------------------------------------------------------

    @property
    def _esize(self):
        return self.get_entry_size()

    def _alloc(self, nbytes):
        try:
            return self._torch.empty(max(1, int(nbytes)), dtype=self._torch.uint8, device=self._device)
        except self._torch.OutOfMemoryError:
            raise OutOfMemoryError

    def _upload(self, dst, host_bytes):
        src = self._torch.from_numpy(np.frombuffer(host_bytes, dtype=np.uint8).copy())
        with self._torch.cuda.stream(self._stream):
            dst[: src.numel()].copy_(src, non_blocking=False)

    def _run(self, launches):
        if not launches:
            return
        arr = _lib.launch_array(launches)
        rc = self._lib.to_run(ctypes.addressof(arr), len(launches), self._entry_type,
                              ctypes.c_void_p(self._stream.cuda_stream))
        if rc == _lib.TO_ERR_OOM:
            raise OutOfMemoryError
        _lib.check(rc, "to_run")

    def synchronize(self):
        _lib.check(self._lib.to_synchronize(ctypes.c_void_p(self._stream.cuda_stream)), "to_synchronize")

    def _budget(self):
        # free device memory plus what torch's caching allocator holds without using it
        free, _total = self._torch.cuda.mem_get_info(self._device)
        cached = self._torch.cuda.memory_reserved(self._device) - self._torch.cuda.memory_allocated(self._device)
        budget = int((free + max(0, cached)) * 0.85)
        if self._mem_limit is not None:
            budget = min(budget, self._mem_limit)
        return budget

    # -- compiled execution ---------------------------------------------------------

    def choose_batch(self, nodes, tables, total_slices):
        """Largest slice batch W (<= max_batch) whose arena fits the device budget; raises
        OutOfMemoryError when even one slice does not fit."""
        budget = self._budget()
        fixed = sum(t.size for t in tables.values()) * self._esize
        W = max(1, min(self._max_batch, total_slices))
        while True:
            peak = schedule.peak_bytes(nodes, tables, W, self._entry_type)
            if peak + fixed <= budget:
                return W
            if W == 1:
                raise OutOfMemoryError
            W = max(1, W // 2)

    def _compile(self, tree, tables, num_groups, total_slices, group_levels=True):
        """Compile and upload a plan; returns an ``Executable``.  ``group_levels=False``
        keeps one launch per join (per-join timing)."""
        nodes = schedule.analyse(tree, tables, self._regroup)
        W = self.choose_batch(nodes, tables, total_slices)
        if self._autotune:
            from . import autotune

            if autotune.tune_plan(self, nodes, W, schedule.TUNED):
                autotune.save(schedule.TUNED)
        prog = schedule.compile_program(tree, tables, num_groups, self._entry_type, W, self._alloc,
                                        regroup=self._regroup, group_levels=group_levels)
        self._upload(prog.table, prog.table_host.tobytes())
        self._upload(prog.coef, prog.coef_host.tobytes())
        self._upload(prog.blob, prog.blob_host.tobytes())
        self._run(prog.build)  # OR / variable leaf tables written on the device
        if self._entry_type == _lib.TO_F32S:
            self._upload(prog.recs, schedule.rec_init_bytes(prog))
        self._upload(prog.params, np.array([0, total_slices, 0, 1], dtype=np.int64).tobytes())
        self._upload(prog.accum, bytes(_lib.Accum()))
        return Executable(self, prog, total_slices)

    def _execute(self, tree, tables, num_groups, total_slices):
        exe = self._compile(tree, tables, num_groups, total_slices)
        try:
            return exe.run()
        finally:
            exe.close()

    def prepare(self, execution_plan, num_slice_limit=None):
        """Compile a sliced plan once; ``Executable.run()`` then contracts all of its
        slices with inputs already resident on the device."""
        total, tree, tables, groups = self._plan_inputs(execution_plan, num_slice_limit)
        return self._compile(tree, tables, groups, total)

    def _plan_inputs(self, execution_plan, num_slice_limit):
        """(slices, slice-free tree, leaf tables, groups) of a plan, as the batched executors
        take them (``jax_apis.py:253-313``): every leaf reordered to [slice row, unsliced
        axes] with its row lookup (``tensor_network.pyx:404-442``, ``tensor.py:95-119``).
        OR and variable leaves are described for the device to write (``device_leaves``);
        other leaves are built on the host."""
        network = execution_plan.network
        groups = [g for g in execution_plan.groups_to_slice if len(g) > 0]
        tree = network.remove_sliced_indices_from(execution_plan.tree, execution_plan.groups_to_slice)
        tables = {}
        if self._device_leaves:
            total, lookups = network.slice_lookups(execution_plan.groups_to_slice)
            for t, lookup in enumerate(lookups):
                spec_table = leaf_spec_table(network[t], lookup, len(groups))
                if spec_table is not None:
                    tables[t] = spec_table
        if len(tables) < len(network):
            total, sequences, _assignments = network.get_tensor_slices(execution_plan.groups_to_slice,
                                                                       self.create_tensor)
            for t, seq in enumerate(sequences):
                if t in tables:
                    continue
                reordered, lookup, shape = seq.reordered_tensor(self)
                reordered = np.asarray(reordered, dtype=np.float64)
                tables[t] = schedule.LeafTable(reordered.reshape(reordered.shape[0], -1), lookup, len(shape))
        if num_slice_limit is not None:
            total = min(total, int(num_slice_limit))
        return total, tree, tables, len(groups)

    def prepare_batch(self, plans):
        """Compile many unsliced plans into grouped schedules (``batch.compile_batch``),
        admitted in waves whose arenas fit the device budget (``mem_limit`` when set): the
        waves share one arena and run one after another.  Raises OutOfMemoryError when one
        instance alone does not fit (slice it first)."""
        from . import batch

        instances = []
        for plan in plans:
            total, tree, tables, groups = self._plan_inputs(plan, None)
            if groups:
                raise ValueError("prepare_batch takes unsliced plans")
            instances.append((tree, tables))
        per = batch.analyse_batch(instances, self._entry_type, regroup=self._regroup)
        waves = batch.admission_waves(per, self._budget())
        if waves is None:
            raise OutOfMemoryError
        arena = self._alloc(max(max(batch.wave_arena_bytes(per, w) for w in waves), schedule.ALIGN))
        progs = []
        for w in waves:
            prog = batch.compile_batch(None, self._entry_type, self._alloc, regroup=self._regroup,
                                       per=[per[i] for i in w], arena=arena)
            self._upload(prog.table, prog.table_host.tobytes())
            self._upload(prog.blob, prog.blob_host.tobytes())
            self._run(prog.build)  # OR / variable leaf tables written on the device
            if self._entry_type == _lib.TO_F32S:
                self._upload(prog.recs, batch.rec_init_bytes(prog))
            progs.append(prog)
        return BatchExecutable(self, progs, arena)

    def contract_batch(self, plans):
        """Counts (floats) of many unsliced plans contracted together."""
        exe = self.prepare_batch(plans)
        try:
            res, _ = exe.run()
        finally:
            exe.close()
        return [r.value for r in res]

    def contract_sliced(self, execution_plan, num_slice_limit=None):
        total, tree, tables, groups = self._plan_inputs(execution_plan, num_slice_limit)
        if total <= 0:
            return 0
        return self._execute(tree, tables, groups, total).value

    def contract(self, network, contraction_tree):
        """Contract one network (sliced tensors keep size-1 axes, which are dropped)."""
        ctx = TreeContext()
        stack, tables = [], {}
        for node in contraction_tree.iterate_postorder():
            if node.is_leaf:
                t = node.tensor_index
                arr = np.asarray(network[t].build(self.create_tensor), dtype=np.float64)
                keep = [e for e, size in zip(node.free_edges, arr.shape) if size != 1]
                tables[t] = schedule.LeafTable(arr.reshape(1, -1), [], len(keep))
                stack.append(ctx.leaf_manual(t, keep))
            else:
                right = stack.pop()
                left = stack.pop()
                stack.append(ctx.join(left, right))
        res = self._execute(ctx.get_tree(stack[0]), tables, 0, 1)
        return np.array(res.value)

    # -- single join ---------------------------------------------------------------

    def _to_device(self, a):
        if isinstance(a, DeviceTensor):
            return a
        arr = np.asarray(a, dtype=np.float64)
        rec = None
        if self._entry_type == _lib.TO_F32S:
            m = float(np.max(np.abs(arr))) if arr.size else 0.0
            e = math.frexp(m)[1] if m > 0 else 0
            host = np.ldexp(arr, -e).astype(np.float32)
            r = _lib.ScaleRec(exp=e, maxbits=int(np.array([1.0], np.float32).view(np.uint32)[0]))
            rec = self._alloc(ctypes.sizeof(r))
            self._upload(rec, bytes(r))
        else:
            host = arr
        data = self._alloc(host.nbytes)
        self._upload(data, np.ascontiguousarray(host).tobytes())
        view = data[: host.nbytes].view(self._torch.float64 if host.dtype == np.float64 else self._torch.float32)
        return DeviceTensor(view, arr.shape, rec, self)

    def tensordot(self, a, b, axes):
        """``numpy.tensordot(a, b, axes)`` on the device; result keeps numpy's axis
        order (left free axes, then right free axes)."""
        A, B = self._to_device(a), self._to_device(b)
        lmap, rmap = list(axes[0]), list(axes[1])
        for (t, m) in ((A, lmap), (B, rmap)):
            if any(d != 2 for d in t.shape):
                raise ValueError("HipAPI tensors must have all dimensions 2")
        g = schedule.join_geometry(A.rank, B.rank, lmap, rmap)
        g["var"] = schedule.choose_variant(g, 1, self._esize)
        out_rank = g["fx"] + g["fy"]
        out_shape = tuple(A.shape[i] for i in range(A.rank) if i not in lmap) + tuple(
            B.shape[i] for i in range(B.rank) if i not in rmap)
        cbuf = self._alloc((1 << out_rank) * self._esize)
        scaled = self._entry_type == _lib.TO_F32S
        crec = self._alloc(ctypes.sizeof(_lib.ScaleRec)) if scaled else None
        X, Y = (A, B) if g["x_is_left"] else (B, A)
        d = _lib.JoinDesc()
        d.x, d.y, d.c = X.data.data_ptr(), Y.data.data_ptr(), cbuf.data_ptr()
        if scaled:
            d.xs, d.ys, d.cs = X.rec.data_ptr(), Y.rec.data_ptr(), crec.data_ptr()
        d.fx, d.fy, d.s, d.shf, d.shy = g["fx"], g["fy"], g["s"], g["shf"], g["shy"]
        d.nbatch = 1
        d.xf_pos, d.xs_pos = schedule._pos_array(g["xf"]), schedule._pos_array(g["xs"])
        d.yf_pos, d.ys_pos = schedule._pos_array(g["yf"]), schedule._pos_array(g["ys"])
        v = g["var"]
        launches = []
        if scaled:
            launches.append(_lib.Launch(op=_lib.TO_OP_SCALE_CLEAR, grid=1, aux0=crec.data_ptr(), iaux0=1))
        keep = [cbuf, crec]
        if v["kind"] == "gemm":
            ops = []
            for need, T, fpos, spos in ((v["perm_x"], X, g["xf"], g["xs"]), (v["perm_y"], Y, g["yf"], g["ys"])):
                if not need:
                    ops.append(T.data.data_ptr())
                    continue
                out_pos = [0] * T.rank
                for t, p in enumerate(spos):
                    out_pos[p] = t
                for i, p in enumerate(fpos):
                    out_pos[p] = len(spos) + i
                dst = self._alloc((1 << T.rank) * self._esize)
                pd, grid, lds, pvar = schedule.perm_desc(T.data.data_ptr(), dst.data_ptr(), T.rank, out_pos, 1,
                                                         1 << T.rank, self._esize)
                pblob = self._alloc(ctypes.sizeof(pd))
                self._upload(pblob, bytes(pd))
                launches.append(_lib.Launch(op=_lib.TO_OP_PERMUTE, variant=pvar, count=1, grid=grid, lds=lds,
                                            descs=pblob.data_ptr()))
                keep += [dst, pblob]
                ops.append(dst.data_ptr())
            gd = _lib.GemmDesc()
            gd.a, gd.b, gd.c = ops[0], ops[1], cbuf.data_ptr()
            gd.ldcm, gd.ldcn = 1 << g["shf"], 1 << g["shy"]
            if scaled:
                gd.as_, gd.bs, gd.cs = X.rec.data_ptr(), Y.rec.data_ptr(), crec.data_ptr()
            gd.mlog, gd.nlog, gd.klog = g["fx"], g["fy"], g["s"]
            gd.nbatch, gd.block0, gd.nblocks = 1, 0, v["nblocks"]
            gblob = self._alloc(ctypes.sizeof(gd))
            self._upload(gblob, bytes(gd))
            launches.append(_lib.Launch(op=_lib.TO_OP_GEMM, variant=v["tile"], count=1, grid=v["nblocks"],
                                        descs=gblob.data_ptr()))
            keep.append(gblob)
        elif v["kind"] == "stream":
            d.nblocks, d.nych, d.nsplit = v["nblocks"], v["nych"], 1
            d.lnych = v["nych"].bit_length() - 1
            d.stage = int(v["stage"])
            blob = self._alloc(ctypes.sizeof(d))
            self._upload(blob, bytes(d))
            variant = schedule.stream_variant(v, g)
            launches.append(_lib.Launch(op=_lib.TO_OP_JOIN_STREAM, variant=variant, count=1,
                                        grid=v["nblocks"], lds=v["lds"], descs=blob.data_ptr()))
            keep.append(blob)
        else:
            part = self._alloc(v["nout"] * v["nsplit"] * 8)  # double partial sums
            d.partial = part.data_ptr()
            d.nsplit, d.nblocks = v["nsplit"], v["nblocks"]
            d2 = _lib.JoinDesc.from_buffer_copy(bytes(d))
            d2.nblocks = v["fin_blocks"]
            blob = self._alloc(2 * ctypes.sizeof(d))
            self._upload(blob, bytes(d) + bytes(d2))
            launches.append(_lib.Launch(op=_lib.TO_OP_JOIN_SPLITK, count=1, grid=v["nblocks"],
                                        descs=blob.data_ptr()))
            launches.append(_lib.Launch(op=_lib.TO_OP_JOIN_FINISH, variant=v["fin_variant"], count=1,
                                        grid=v["fin_blocks"], descs=blob.data_ptr() + ctypes.sizeof(d)))
            keep += [blob, part]
        self._run(launches)
        self.synchronize()
        tdtype = self._torch.float64 if not scaled else self._torch.float32
        view = cbuf[: (1 << out_rank) * self._esize].view(tdtype)
        return DeviceTensor(view, out_shape, crec, self)

    def permute(self, tensor, perm):
        """Device ``numpy.transpose`` of an all-dimension-2 tensor (returns numpy)."""
        A = self._to_device(tensor)
        out = self._alloc((1 << A.rank) * self._esize)
        p = (ctypes.c_int32 * max(1, len(perm)))(*perm)
        rc = self._lib.to_permute(ctypes.c_void_p(A.data.data_ptr()), ctypes.c_void_p(out.data_ptr()), A.rank,
                                  ctypes.addressof(p), self._esize, 1, ctypes.c_void_p(self._stream.cuda_stream))
        _lib.check(rc, "to_permute")
        self.synchronize()
        tdtype = self._torch.float64 if self._esize == 8 else self._torch.float32
        view = out[: (1 << A.rank) * self._esize].view(tdtype)
        return DeviceTensor(view, tuple(A.shape[i] for i in perm), A.rec, self).numpy()


class HipF32API(HipAPI):
    """The FP32-scaled executor as its own library choice."""

    def __init__(self, device=0):
        super().__init__(device)
        self._entry_type = _lib.TO_F32S


HIP_APIS = {"hip": HipAPI, "hip-f32": HipF32API}
