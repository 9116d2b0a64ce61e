"""to_tensordot_noalloc (the allocation-free tensordot: split-K partial sums in the
library's file-scope scratch array) against to_tensordot (partial sums in a workspace the
caller allocates): the same kernels, so the same bits, on two calls in a row, the second
with a smaller input; and a call needing more than the scratch array is refused."""

import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


# (rank a, rank b, shared): split-K forms first large, then smaller (both use the scratch),
# then a streamed join (no scratch)
CALLS = [(20, 20, 20), (16, 16, 14), (14, 12, 10), (12, 10, 6)]


@pytest.mark.parametrize("et", ["f64", "f32s"])
def test_noalloc_matches_workspace_entry(et):
    import torch

    from ctensordot import c_tensordot, workspace_bytes

    from tensororder_amd import _lib

    lib = _lib.load()
    stream = torch.cuda.Stream()
    entry_type = _lib.TO_F64 if et == "f64" else _lib.TO_F32S
    rng = np.random.default_rng(23)
    sizes = []
    for ra, rb, s in CALLS:
        a = rng.uniform(0.5, 2.0, (2,) * ra)
        b = rng.uniform(0.5, 2.0, (2,) * rb)
        if et == "f32s":
            a = a.astype(np.float32).astype(np.float64)
            b = b.astype(np.float32).astype(np.float64)
        la = [int(x) for x in rng.permutation(ra)[:s]]
        lb = [int(x) for x in rng.permutation(rb)[:s]]
        sizes.append(workspace_bytes(lib, ra, rb, la, lb, entry_type))
        _say("case noalloc", et, ra, rb, s, "partial bytes", sizes[-1])
        want = c_tensordot(lib, stream, a, b, (la, lb), entry_type, "to_tensordot")
        got = c_tensordot(lib, stream, a, b, (la, lb), entry_type, "to_tensordot_noalloc")
        np.testing.assert_array_equal(got, want)
        np.testing.assert_allclose(got, np.tensordot(a, b, (la, lb)), rtol=1e-12 if et == "f64" else 1e-6)
    assert sizes[0] > sizes[1] > sizes[2] > 0 and sizes[3] == 0


def test_noalloc_refuses_more_than_its_scratch():
    import torch

    from ctensordot import _Axes

    from tensororder_amd import _lib

    lib = _lib.load()
    torch.cuda.init()
    # 2^21 outputs over 13 shared axes: 2^21 partial sums (16 MiB) exceed the 8 MiB array
    ax = _Axes(list(range(13)), list(range(13)))
    rc = lib.to_tensordot_noalloc(ctypes.c_void_p(256), ctypes.c_void_p(256), ctypes.c_void_p(256), 24, 23,
                                  ax[0], ax[1], 13, _lib.TO_F64, None, None, None, None)
    assert rc == _lib.TO_ERR_INVALID
    assert b"scratch" in lib.to_last_error()
