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
    sizes = [ctypes.c_int64() for _ in range(8)]
    lib.to_struct_sizes(*[ctypes.byref(s) for s in sizes])
    assert [s.value for s in sizes] == [ctypes.sizeof(t) for t in (
        _lib.JoinDesc, _lib.LeafDesc, _lib.ScaleRec, _lib.Accum, _lib.Launch, _lib.PermDesc, _lib.GemmDesc,
        _lib.RootDesc)]


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
