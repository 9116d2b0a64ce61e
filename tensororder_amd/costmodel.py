"""Learned plan-cost model: per-join GPU cost from features, and plan choice by it.

The reference chooses a plan by the smallest max rank (``planning.py:159-162``), stops
planning by ``performance_factor * FLOPs`` (``planning.py:168-172``) and slices with the
``greedy_mem`` cost model (``contraction_tree.pyx:382-496``).  Here an FP32 MLP
(16 -> 64 -> 64 -> 2, trained and evaluated by the HIP kernels of
``include/tensororder_costmodel.h``) predicts log device time and log memory footprint of
every launch group of a compiled plan, from the executor's own timed records.  The
predicted plan time is the sum over the prologue's joins plus the number of waves times
the sum over a wave's joins; the candidate with the smallest prediction is chosen.  The
reference rule remains the fallback, and the choice never changes the computed count --
every candidate is an exact plan of the same network.
"""

import ctypes
import math
import os

import numpy as np

from . import _lib, schedule

FEATURES = [
    "rank_x", "rank_y", "rank_out", "shared", "free_x", "free_y", "log2_flops", "log2_bytes",
    "log2_batch", "stream", "splitk", "gemm", "entry_bytes_over_8", "permutes", "log2_outputs",
    "variant",
]
TARGETS = ["log_time_us", "log_footprint_bytes"]

# parameters trained by ``python -m tensororder_amd.train_costmodel`` on the executor's own
# timed records (see the summary json beside them)
DEFAULT_PARAMS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "costmodel_params.npz")


def join_features(n, nodes, W, esize):
    """Raw feature vector (len 16) of one join of an analysed node list."""
    g, v = n.geom, n.geom["var"]
    nb = W if n.batched else 1
    X, Y = (nodes[n.left], nodes[n.right]) if g["x_is_left"] else (nodes[n.right], nodes[n.left])
    b, f = schedule.join_traffic(n, nodes, W, esize)
    kind = v["kind"]
    perms = (int(v.get("perm_x", False)) + int(v.get("perm_y", False))) if kind in ("gemm", "splitk") else 0
    variant = {"stream": 10 if v.get("rows") else math.log2(v.get("yc", 1)) + (5 if v.get("vec") else 0),
               "gemm": v.get("tile", 0) & 7}.get(kind, 0)
    return [X.rank, Y.rank, n.rank, g["s"], g["fx"], g["fy"], math.log2(max(f, 1)), math.log2(max(b, 1)),
            math.log2(nb), kind == "stream", kind == "splitk", kind == "gemm", esize / 8.0, perms,
            g["fx"] + g["fy"] + math.log2(nb), variant]


def join_footprint(n, nodes, W, esize):
    """Bytes live while a join runs: operands, output and scratch."""
    nb = W if n.batched else 1
    total = 0
    for ch in (nodes[n.left], nodes[n.right]):
        total += (1 << ch.rank) * (nb if ch.batched else 1)
    total += (1 << n.rank) * nb
    v = n.geom["var"]
    if v["kind"] == "splitk":
        total += v["nout"] * v["nsplit"]
    if v["kind"] in ("gemm", "splitk"):
        for key, ch in (("perm_x", nodes[n.left]), ("perm_y", nodes[n.right])):
            if v.get(key):
                total += (1 << ch.rank) * (nb if ch.batched else 1)
    return total * esize


def records_from_timing(prog, infos, ms, W, esize):
    """(features, targets) of every join of one timed step: a join's time is the sum of its
    launches (permutes + product, or split-K + finish)."""
    per_node = {}
    for (kind, nidx), t in zip(infos, ms):
        if nidx is None or kind in ("leaf", "root", "clear", "advance", "begin"):
            continue
        per_node[nidx] = per_node.get(nidx, 0.0) + float(t)
    X, Y = [], []
    for nidx, t in per_node.items():
        n = prog.nodes[nidx]
        X.append(join_features(n, prog.nodes, W, esize))
        Y.append([math.log(max(t, 1e-6) * 1e3), math.log(join_footprint(n, prog.nodes, W, esize))])
    return np.asarray(X, dtype=np.float32), np.asarray(Y, dtype=np.float32)


def plan_flops(nodes, tables, esize, entry_type):
    """(flops per slice of the batched joins, flops of the slice-invariant joins) of an
    analysed plan."""
    schedule.peak_bytes(nodes, tables, 1, entry_type)
    per_slice = const = 0.0
    for n in nodes:
        if n.is_leaf:
            continue
        _, f = schedule.join_traffic(n, nodes, 1, esize)
        if n.batched:
            per_slice += f
        else:
            const += f
    return per_slice, const


def plan_launch_features(nodes, W, esize, nwaves):
    """Features of every join of an analysed plan and how many times it runs per contraction."""
    feats, mult = [], []
    for n in nodes:
        if n.is_leaf:
            continue
        feats.append(join_features(n, nodes, W, esize))
        mult.append(nwaves if n.batched else 1)
    return np.asarray(feats, dtype=np.float32), np.asarray(mult, dtype=np.float64)


class CostModel:
    """The MLP with device-resident parameters; training and inference run in HIP."""

    def __init__(self, api, seed=0):
        self.api = api
        self.lib = api._lib
        torch = api._torch
        self.torch = torch
        rng = np.random.default_rng(seed)
        parts = []
        for dout, din in ((64, 16), (64, 64), (2, 64)):
            parts.append(rng.standard_normal((dout, din)) * math.sqrt(2.0 / din))
            parts.append(np.zeros(dout))
        host = np.concatenate([p.ravel() for p in parts]).astype(np.float32)
        assert host.size == _lib.MLP_NPARAMS
        dev = api._device
        self.params = torch.from_numpy(host).to(dev)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grad = torch.zeros_like(self.params)
        self.step = 0
        self.x_mean = np.zeros(16, np.float32)
        self.x_std = np.ones(16, np.float32)
        self.y_mean = np.zeros(2, np.float32)
        self.y_std = np.ones(2, np.float32)

    def _ptr(self, t):
        return ctypes.c_void_p(t.data_ptr())

    def _stream(self):
        return ctypes.c_void_p(self.api._stream.cuda_stream)

    def _buffers(self, n):
        torch, dev = self.torch, self.api._device
        return (torch.empty((n, 64), dtype=torch.float32, device=dev),
                torch.empty((n, 64), dtype=torch.float32, device=dev),
                torch.empty((n, 2), dtype=torch.float32, device=dev))

    def forward_raw(self, xdev):
        """Device forward pass on normalised inputs; returns (h1, h2, y) tensors."""
        n = xdev.shape[0]
        h1, h2, y = self._buffers(n)
        with self.torch.cuda.stream(self.api._stream):
            rc = self.lib.to_mlp_forward(self._ptr(self.params), self._ptr(xdev), n, self._ptr(h1), self._ptr(h2),
                                         self._ptr(y), self._stream())
        _lib.check(rc, "to_mlp_forward")
        return h1, h2, y

    def gradient(self, xdev, tdev):
        """(loss, gradient) on device data (normalised)."""
        n = xdev.shape[0]
        h1, h2, y = self.forward_raw(xdev)
        dh1, dh2, dy = self._buffers(n)
        loss = self.torch.zeros(1, dtype=self.torch.float32, device=self.api._device)
        rc = self.lib.to_mlp_backward(self._ptr(self.params), self._ptr(xdev), self._ptr(h1), self._ptr(h2),
                                      self._ptr(y), self._ptr(tdev), n, self._ptr(dy), self._ptr(dh2),
                                      self._ptr(dh1), self._ptr(self.grad), self._ptr(loss), self._stream())
        _lib.check(rc, "to_mlp_backward")
        self.api.synchronize()
        return float(loss.cpu()[0]), self.grad

    def adam_step(self, lr=3e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        rc = self.lib.to_mlp_adam(self._ptr(self.params), self._ptr(self.grad), self._ptr(self.m), self._ptr(self.v),
                                  self.params.numel(), lr, b1, b2, eps, self.step, self._stream())
        _lib.check(rc, "to_mlp_adam")

    def fit(self, X, Y, epochs=300, batch=2048, lr=3e-3, seed=0):
        """Train on (features, targets); returns the per-epoch mean loss (normalised)."""
        X = np.asarray(X, np.float32)
        Y = np.asarray(Y, np.float32)
        self.x_mean, self.x_std = X.mean(0), X.std(0) + 1e-6
        self.y_mean, self.y_std = Y.mean(0), Y.std(0) + 1e-6
        torch, dev = self.torch, self.api._device
        xd = torch.from_numpy((X - self.x_mean) / self.x_std).to(dev)
        yd = torch.from_numpy((Y - self.y_mean) / self.y_std).to(dev)
        rng = np.random.default_rng(seed)
        history = []
        n = X.shape[0]
        for _ in range(epochs):
            order = torch.from_numpy(rng.permutation(n)).to(dev)
            tot, cnt = 0.0, 0
            for s in range(0, n, batch):
                idx = order[s:s + batch]
                xb, yb = xd.index_select(0, idx).contiguous(), yd.index_select(0, idx).contiguous()
                loss, _ = self.gradient(xb, yb)
                self.adam_step(lr)
                tot += loss * len(idx)
                cnt += len(idx)
            history.append(tot / cnt)
        return history

    def predict(self, X):
        """(log time us, log footprint bytes) per row, in the targets' units."""
        X = np.asarray(X, np.float32)
        if X.shape[0] == 0:
            return np.zeros((0, 2), np.float32)
        xd = self.torch.from_numpy((X - self.x_mean) / self.x_std).to(self.api._device)
        _, _, y = self.forward_raw(xd)
        self.api.synchronize()
        return y.cpu().numpy() * self.y_std + self.y_mean

    def state(self):
        return dict(params=self.params.cpu().numpy(), x_mean=self.x_mean, x_std=self.x_std,
                    y_mean=self.y_mean, y_std=self.y_std)

    def load_state(self, st):
        self.params.copy_(self.torch.from_numpy(np.asarray(st["params"], np.float32)))
        self.x_mean, self.x_std = np.asarray(st["x_mean"]), np.asarray(st["x_std"])
        self.y_mean, self.y_std = np.asarray(st["y_mean"]), np.asarray(st["y_std"])

    def save(self, path):
        np.savez(path, **self.state())

    def load(self, path):
        with np.load(path, allow_pickle=False) as z:
            self.load_state({k: z[k] for k in z.files})


def predicted_seconds(model, api, plan):
    """Predicted device seconds of one full contraction of ``plan`` by ``api``."""
    total, tree, tables, groups = api._plan_inputs(plan, None)
    nodes = schedule.analyse(tree, tables, api._regroup)
    W = api.choose_batch(nodes, tables, total)
    esize = api.get_entry_size()
    schedule.peak_bytes(nodes, tables, W, api._entry_type)
    nwaves = -(-total // W)
    feats, mult = plan_launch_features(nodes, W, esize, nwaves)
    if feats.shape[0] == 0:
        return 0.0
    pred = model.predict(feats)
    return float(np.sum(np.exp(pred[:, 0]) * mult) * 1e-6)


def reference_choice(plans):
    """The reference rule: smallest max rank, earliest on ties (planning.py:159-162)."""
    best = 0
    for i, p in enumerate(plans):
        if p.maxrank < plans[best].maxrank:
            best = i
    return best


def flop_choice(plans):
    """A FLOP rule: the candidate of smallest estimated FLOPs (the reference's cost model,
    contraction_tree.pyx:449-458, times the slice count), earliest on ties."""
    best = 0
    for i, p in enumerate(plans):
        if p.total_FLOPs < plans[best].total_FLOPs:
            best = i
    return best


def learned_choice(model, api, plans):
    """Index of the candidate with the smallest predicted device time, and all predictions.
    A candidate that does not fit the device even one slice at a time is never chosen."""
    from .api import OutOfMemoryError

    preds = []
    for p in plans:
        try:
            preds.append(predicted_seconds(model, api, p))
        except OutOfMemoryError:
            preds.append(math.inf)
    return int(np.argmin(preds)), preds


def load_default(api):
    """The trained model of ``DEFAULT_PARAMS`` on ``api``'s device."""
    model = CostModel(api)
    model.load(DEFAULT_PARAMS)
    return model
