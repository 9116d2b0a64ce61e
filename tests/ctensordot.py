"""Test helper: numpy.tensordot through the library's C entry points (to_tensordot with a
torch workspace, or another entry point of the same signature without the workspace
arguments), on all-dimension-2 tensors, in float64 or FP32 with scale records."""

import ctypes
import math

import numpy as np

from tensororder_amd import _lib


class _Axes:
    def __init__(self, *lists):
        self.arrays = [(ctypes.c_int32 * max(1, len(v)))(*v) for v in lists]

    def __getitem__(self, i):
        return ctypes.addressof(self.arrays[i])


def _rec(torch, dev, arr):
    """(device float32 data scaled to max |entry| <= 1, device to_scale_rec) of a float64 array."""
    m = float(np.max(np.abs(arr))) if arr.size else 0.0
    e = math.frexp(m)[1] if m > 0 else 0
    data = np.ldexp(arr, -e).astype(np.float32)
    maxbits = int(np.float32(np.max(np.abs(data)) if data.size else 0.0).view(np.uint32))
    rec = np.zeros(2, dtype=np.int64)
    rec[0] = e
    rec[1] = maxbits
    return torch.from_numpy(data.ravel().copy()).to(dev), torch.from_numpy(rec).to(dev)


def workspace_bytes(lib, ra, rb, la, lb, entry_type):
    ax = _Axes(la, lb)
    n = ctypes.c_int64()
    _lib.check(lib.to_tensordot_workspace(ra, rb, ax[0], ax[1], len(la), entry_type, ctypes.byref(n)),
               "to_tensordot_workspace")
    return n.value


def c_tensordot(lib, stream, a, b, axes, entry_type, entry="to_tensordot"):
    """Run one C tensordot; returns the result as float64 numpy (scaled back in FP32 mode)."""
    import torch

    dev = torch.device("cuda", 0)
    la, lb = [int(x) for x in axes[0]], [int(x) for x in axes[1]]
    ra, rb = a.ndim, b.ndim
    out_rank = ra + rb - 2 * len(la)
    ax = _Axes(la, lb)
    scaled = entry_type == _lib.TO_F32S
    if scaled:
        ta, ra_rec = _rec(torch, dev, np.asarray(a, np.float64))
        tb, rb_rec = _rec(torch, dev, np.asarray(b, np.float64))
        tc = torch.full((1 << out_rank,), float("nan"), dtype=torch.float32, device=dev)
        rc_rec = torch.full((2,), -1, dtype=torch.int64, device=dev)
        recs = (ra_rec.data_ptr(), rb_rec.data_ptr(), rc_rec.data_ptr())
    else:
        ta = torch.from_numpy(np.ascontiguousarray(a, np.float64).ravel().copy()).to(dev)
        tb = torch.from_numpy(np.ascontiguousarray(b, np.float64).ravel().copy()).to(dev)
        tc = torch.full((1 << out_rank,), float("nan"), dtype=torch.float64, device=dev)
        recs = (None, None, None)
    torch.cuda.synchronize()
    args = [ctypes.c_void_p(ta.data_ptr()), ctypes.c_void_p(tb.data_ptr()), ctypes.c_void_p(tc.data_ptr()),
            ra, rb, ax[0], ax[1], len(la), entry_type] + [ctypes.c_void_p(r) if r else None for r in recs]
    if entry == "to_tensordot":
        need = workspace_bytes(lib, ra, rb, la, lb, entry_type)
        ws = torch.empty(max(need, 8), dtype=torch.uint8, device=dev)
        args += [ctypes.c_void_p(ws.data_ptr()), need]
    args.append(ctypes.c_void_p(stream.cuda_stream))
    _lib.check(getattr(lib, entry)(*args), entry)
    stream.synchronize()
    out = tc.cpu().numpy().astype(np.float64)
    if scaled:
        exp = int(rc_rec.cpu().numpy()[0])
        maxbits = int(rc_rec.cpu().numpy()[1]) & 0xffffffff
        assert maxbits == int(np.float32(np.max(np.abs(tc.cpu().numpy()))).view(np.uint32))
        out = np.ldexp(out, exp)
    shape = tuple(2 for _ in range(out_rank))
    return out.reshape(shape) if shape else out.reshape(())
