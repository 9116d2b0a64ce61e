"""The C-ABI library loads and exports every entry point include/tensororder_hip.h
declares, with record layouts matching the ctypes mirror (no GPU needed)."""

import ctypes
import os
import re

from tensororder_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    import glob

    text = "".join(open(p).read() for p in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))))
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(to_\w+)\s*\(", text, flags=re.M)))


def test_exports_every_declared_symbol():
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 10
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(_lib.EXPORTED) == names


def test_layouts_and_version():
    lib = _lib.load()
    assert lib.to_version() == _lib.ABI_VERSION
    sizes = [ctypes.c_int64() for _ in range(9)]
    lib.to_struct_sizes(*[ctypes.byref(s) for s in sizes])
    assert [s.value for s in sizes] == [ctypes.sizeof(t) for t in (
        _lib.JoinDesc, _lib.LeafDesc, _lib.ScaleRec, _lib.Accum, _lib.Launch, _lib.PermDesc, _lib.GemmDesc,
        _lib.RootDesc, _lib.LeafBuildDesc)]


def test_invalid_records_are_rejected_without_launching():
    lib = _lib.load()
    bad = _lib.Launch(op=99, grid=1)
    arr = _lib.launch_array([bad])
    assert lib.to_run(ctypes.addressof(arr), 1, _lib.TO_F64, None) == _lib.TO_ERR_INVALID
    assert b"unknown" in lib.to_last_error()
    neg = _lib.Launch(op=_lib.TO_OP_JOIN_STREAM, grid=0)
    arr = _lib.launch_array([neg])
    assert lib.to_run(ctypes.addressof(arr), 1, _lib.TO_F64, None) == _lib.TO_ERR_INVALID
    assert lib.to_run(ctypes.addressof(arr), 1, 7, None) == _lib.TO_ERR_INVALID
    # a one-wave join launch whose grid holds fewer waves than descriptors
    waves = _lib.Launch(op=_lib.TO_OP_JOIN_STREAM, variant=_lib.TO_STREAM_WAVE, count=9, grid=2, descs=256)
    arr = _lib.launch_array([waves])
    assert lib.to_run(ctypes.addressof(arr), 1, _lib.TO_F64, None) == _lib.TO_ERR_INVALID
    assert b"wave" in lib.to_last_error()
    perm = (ctypes.c_int32 * 2)(0, 0)
    assert lib.to_permute(ctypes.c_void_p(8), ctypes.c_void_p(8), 2, ctypes.addressof(perm), 8, 1,
                          None) == _lib.TO_ERR_INVALID


def test_stream_any_descriptors_name_known_forms():
    # a TO_STREAM_ANY launch runs each descriptor's own form; the host rejects a
    # descriptor naming a form the kernel does not have instead of launching it
    import pytest

    from tensororder_amd import schedule

    for kvar in schedule.STREAM_FORMS:
        schedule.check_stream_form(kvar)
    for bad in (0, 3, 32, _lib.TO_STREAM_ANY, _lib.TO_STREAM_VEC, _lib.TO_STREAM_ROWS | 4):
        with pytest.raises(ValueError):
            schedule.check_stream_form(bad)


class _Axes:
    """int32 axis arrays kept alive for the duration of a call."""

    def __init__(self, *lists):
        self.arrays = [(ctypes.c_int32 * max(1, len(v)))(*v) for v in lists]

    def __getitem__(self, i):
        return ctypes.addressof(self.arrays[i])


def test_tensordot_workspace_sizes_and_argument_checks():
    # the C tensordot's plan is host logic: workspace sizes and argument errors need no GPU
    lib = _lib.load()
    nbytes = ctypes.c_int64(-1)

    def ws(ra, rb, la, lb, et):
        ax = _Axes(la, lb)
        rc = lib.to_tensordot_workspace(ra, rb, ax[0], ax[1], len(la), et, ctypes.byref(nbytes))
        return rc, nbytes.value

    # a streamed join needs no workspace
    assert ws(12, 10, [1, 4], [0, 9], _lib.TO_F64) == (_lib.TO_OK, 0)
    # a dot product of two rank-20 tensors: split-K, 4096 double partial sums of one output
    assert ws(20, 20, list(range(20)), list(range(20)), _lib.TO_F32S) == (_lib.TO_OK, 4096 * 8)
    # 16 outputs, 14 shared axes: 64 splits each
    assert ws(16, 16, list(range(2, 16)), list(range(14)), _lib.TO_F64) == (_lib.TO_OK, 16 * 64 * 8)
    for ra, rb, la, lb, et in ((3, 3, [0, 0], [1, 2], _lib.TO_F64),   # repeated axis
                               (3, 3, [3], [0], _lib.TO_F64),         # axis out of range
                               (31, 2, [0], [0], _lib.TO_F64),        # rank above 30
                               (20, 20, [0], [0], _lib.TO_F64),       # output rank 38
                               (3, 3, [0], [0], 7)):                  # entry type
        assert ws(ra, rb, la, lb, et)[0] == _lib.TO_ERR_INVALID
    # a split-K tensordot without its workspace is rejected before any launch
    ax = _Axes(list(range(20)), list(range(20)))
    rc = lib.to_tensordot(ctypes.c_void_p(256), ctypes.c_void_p(256), ctypes.c_void_p(256), 20, 20, ax[0], ax[1],
                          20, _lib.TO_F64, None, None, None, None, 0, None)
    assert rc == _lib.TO_ERR_INVALID and b"workspace" in lib.to_last_error()
