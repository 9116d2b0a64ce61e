"""ctypes binding of ``libtensororder_hip.so`` (``include/tensororder_hip.h``).

The library is built in-tree by ``tensororder_amd.build``.  Loading fails loudly when it
is missing: there is no CPU fallback on the product path.
"""

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtensororder_hip.so")

TO_OK, TO_ERR_OOM, TO_ERR_INVALID, TO_ERR_HIP = 0, 1, 2, 3
TO_F64, TO_F32S = 0, 1
(
    TO_OP_LEAF_GATHER,
    TO_OP_JOIN_STREAM,
    TO_OP_JOIN_SPLITK,
    TO_OP_JOIN_FINISH,
    TO_OP_SCALE_CLEAR,
    TO_OP_ROOT_REDUCE,
    TO_OP_WAVE_ADVANCE,
    TO_OP_STEP_BEGIN,
    TO_OP_PERMUTE,
    TO_OP_GEMM,
    TO_OP_ROOT_STORE,
    TO_OP_BATCH_BEGIN,
    TO_OP_LEAF_BUILD,
) = range(1, 14)
TO_LEAF_VAR, TO_LEAF_OR = 0, 1  # leaf build kinds
TO_STREAM_VEC = 0x100  # stream variant bit: the 16-byte form
TO_STREAM_ROWS = 0x200  # stream variant: the row form
TO_STREAM_ANY = 0x400  # stream launch variant: mixed joins, each by its descriptor's kvar
TO_STREAM_WAVE = 0x800  # stream launch variant: one wave per (small) join, four per block
TO_LAUNCH_SIDE = 0x1  # Launch.flags: captured on the side stream
TO_LAUNCH_JOIN = 0x2  # Launch.flags: join the side stream into the main one and fork it again first
MAX_BITS = 32
ABI_VERSION = 6

c_i8x32 = ctypes.c_int8 * MAX_BITS


class JoinDesc(ctypes.Structure):
    _fields_ = [
        ("x", ctypes.c_uint64), ("y", ctypes.c_uint64), ("c", ctypes.c_uint64),
        ("xb", ctypes.c_int64), ("yb", ctypes.c_int64), ("cb", ctypes.c_int64),
        ("xs", ctypes.c_uint64), ("ys", ctypes.c_uint64), ("cs", ctypes.c_uint64),
        ("xsb", ctypes.c_int64), ("ysb", ctypes.c_int64),
        ("partial", ctypes.c_uint64),
        ("fx", ctypes.c_int32), ("fy", ctypes.c_int32), ("s", ctypes.c_int32),
        ("shf", ctypes.c_int32), ("shy", ctypes.c_int32),
        ("nbatch", ctypes.c_int32),
        ("block0", ctypes.c_int32), ("nblocks", ctypes.c_int32),
        ("nych", ctypes.c_int32), ("nsplit", ctypes.c_int32),
        ("lnych", ctypes.c_int32), ("stage", ctypes.c_int32), ("kvar", ctypes.c_int32),
        ("xf_pos", c_i8x32), ("xs_pos", c_i8x32), ("yf_pos", c_i8x32), ("ys_pos", c_i8x32),
    ]


class LeafDesc(ctypes.Structure):
    _fields_ = [
        ("src", ctypes.c_uint64), ("dst", ctypes.c_uint64), ("coef", ctypes.c_uint64),
        ("rec", ctypes.c_uint64), ("exp0", ctypes.c_int64),
        ("rank", ctypes.c_int32), ("ngroups", ctypes.c_int32),
        ("nbatch", ctypes.c_int32), ("block0", ctypes.c_int32), ("nblocks", ctypes.c_int32),
        ("pad", ctypes.c_int32),
    ]


class LeafBuildDesc(ctypes.Structure):
    _fields_ = [
        ("dst", ctypes.c_uint64), ("val0", ctypes.c_double), ("val1", ctypes.c_double),
        ("negmask", ctypes.c_int64),
        ("kind", ctypes.c_int32), ("rank", ctypes.c_int32), ("out_bit", ctypes.c_int32),
        ("block0", ctypes.c_int32), ("nblocks", ctypes.c_int32), ("pad", ctypes.c_int32),
        ("pos", ctypes.c_int8 * 32),
    ]


class ScaleRec(ctypes.Structure):
    _fields_ = [("exp", ctypes.c_int64), ("maxbits", ctypes.c_uint32), ("pad", ctypes.c_uint32)]


class Accum(ctypes.Structure):
    _fields_ = [("mant", ctypes.c_double), ("exp", ctypes.c_int64),
                ("count", ctypes.c_int64), ("pad", ctypes.c_int64)]


c_i8x16 = ctypes.c_int8 * 16


class PermDesc(ctypes.Structure):
    _fields_ = [
        ("in_", ctypes.c_uint64), ("out", ctypes.c_uint64),
        ("inb", ctypes.c_int64), ("outb", ctypes.c_int64),
        ("rank", ctypes.c_int32), ("nu", ctypes.c_int32), ("nrest", ctypes.c_int32), ("nbatch", ctypes.c_int32),
        ("block0", ctypes.c_int32), ("nblocks", ctypes.c_int32),
        ("u_in", c_i8x16), ("v_out", c_i8x16), ("v2u", c_i8x16),
        ("rest_in", c_i8x32), ("rest_out", c_i8x32),
    ]


class GemmDesc(ctypes.Structure):
    _fields_ = [
        ("a", ctypes.c_uint64), ("b", ctypes.c_uint64), ("c", ctypes.c_uint64),
        ("ab", ctypes.c_int64), ("bb", ctypes.c_int64), ("cb", ctypes.c_int64),
        ("ldcm", ctypes.c_int64), ("ldcn", ctypes.c_int64),
        ("as_", ctypes.c_uint64), ("bs", ctypes.c_uint64), ("cs", ctypes.c_uint64),
        ("asb", ctypes.c_int64), ("bsb", ctypes.c_int64),
        ("mlog", ctypes.c_int32), ("nlog", ctypes.c_int32), ("klog", ctypes.c_int32),
        ("nbatch", ctypes.c_int32), ("block0", ctypes.c_int32), ("nblocks", ctypes.c_int32),
        ("lsplit", ctypes.c_int32), ("pad1", ctypes.c_int32),
    ]


class RootDesc(ctypes.Structure):
    _fields_ = [("root", ctypes.c_uint64), ("rec", ctypes.c_uint64), ("acc", ctypes.c_uint64),
                ("block0", ctypes.c_int32), ("nblocks", ctypes.c_int32)]


class Launch(ctypes.Structure):
    _fields_ = [
        ("op", ctypes.c_int32), ("variant", ctypes.c_int32), ("count", ctypes.c_int32),
        ("grid", ctypes.c_int32), ("lds", ctypes.c_int32), ("flags", ctypes.c_int32),
        ("descs", ctypes.c_uint64),
        ("aux0", ctypes.c_uint64), ("aux1", ctypes.c_uint64),
        ("aux2", ctypes.c_uint64), ("aux3", ctypes.c_uint64),
        ("iaux0", ctypes.c_int64), ("iaux1", ctypes.c_int64),
    ]


EXPORTED = [
    "to_version", "to_struct_sizes", "to_last_error", "to_device_count", "to_set_device",
    "to_run", "to_run_timed", "to_graph_create", "to_graph_launch", "to_graph_destroy", "to_permute",
    "to_wall_clock_khz", "to_synchronize", "to_mlp_forward", "to_mlp_backward", "to_mlp_adam",
    "to_tensordot_workspace", "to_tensordot", "to_tensordot_noalloc",
]
MLP_IN, MLP_H1, MLP_H2, MLP_OUT = 16, 64, 64, 2
MLP_NPARAMS = 64 * 16 + 64 + 64 * 64 + 64 + 2 * 64 + 2

_lib = None


class HipError(RuntimeError):
    pass


def load(path=LIB_PATH):
    """Load and check the library (cached)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(
            "libtensororder_hip.so is not built; run `python -m tensororder_amd.build`"
        )
    lib = ctypes.CDLL(path)
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
    p64 = ctypes.POINTER(ctypes.c_int64)
    lib.to_version.restype = i32
    lib.to_struct_sizes.argtypes = [p64] * 9
    lib.to_wall_clock_khz.argtypes = [ctypes.POINTER(ctypes.c_int)]
    lib.to_last_error.restype = ctypes.c_char_p
    lib.to_device_count.argtypes = [ctypes.POINTER(ctypes.c_int)]
    lib.to_set_device.argtypes = [i32]
    lib.to_run.argtypes = [vp, i32, i32, vp]
    lib.to_run_timed.argtypes = [vp, i32, i32, vp, vp]
    lib.to_graph_create.argtypes = [vp, i32, i32, vp, ctypes.POINTER(vp)]
    lib.to_graph_launch.argtypes = [vp, vp]
    lib.to_graph_destroy.argtypes = [vp]
    lib.to_permute.argtypes = [vp, vp, i32, vp, i32, i64, vp]
    lib.to_synchronize.argtypes = [vp]
    lib.to_tensordot_workspace.argtypes = [i32, i32, vp, vp, i32, i32, p64]
    lib.to_tensordot.argtypes = [vp, vp, vp, i32, i32, vp, vp, i32, i32, vp, vp, vp, vp, i64, vp]
    lib.to_tensordot_noalloc.argtypes = [vp, vp, vp, i32, i32, vp, vp, i32, i32, vp, vp, vp, vp]
    lib.to_mlp_forward.argtypes = [vp, vp, i64, vp, vp, vp, vp]
    lib.to_mlp_backward.argtypes = [vp, vp, vp, vp, vp, vp, i64, vp, vp, vp, vp, vp, vp]
    lib.to_mlp_adam.argtypes = [vp, vp, vp, vp, i64, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                ctypes.c_float, i64, vp]
    if lib.to_version() != ABI_VERSION:
        raise ImportError("libtensororder_hip.so ABI version mismatch; rebuild it")
    sizes = [ctypes.c_int64() for _ in range(9)]
    lib.to_struct_sizes(*[ctypes.byref(s) for s in sizes])
    want = [ctypes.sizeof(t) for t in (JoinDesc, LeafDesc, ScaleRec, Accum, Launch, PermDesc, GemmDesc, RootDesc,
                                       LeafBuildDesc)]
    if [s.value for s in sizes] != want:
        raise ImportError("record layout mismatch between library and binding: %s vs %s"
                          % ([s.value for s in sizes], want))
    _lib = lib
    return lib


def check(rc, what):
    """Raise on a non-OK status; OOM is mapped by the caller."""
    if rc != TO_OK:
        msg = load().to_last_error().decode(errors="replace")
        raise HipError("%s failed (status %d): %s" % (what, rc, msg))


def launch_array(records):
    arr = (Launch * len(records))(*records)
    return arr
