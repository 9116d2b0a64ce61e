"""Build the HIP executor library in-tree for gfx950.

``python -m tensororder_amd.build`` compiles ``csrc/tensororder_hip.hip`` and
``csrc/costmodel.hip`` into
``tensororder_amd/libtensororder_hip.so`` with hipcc.  Packed-FP32 instructions are
disabled for all device code (-packed-fp32-ops).
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "tensororder_hip.hip")
SOURCES = [SRC, os.path.join(HERE, "csrc", "costmodel.hip")]
OUT = os.path.join(HERE, "libtensororder_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TENSORORDER_ARCH", "gfx950")


def hipcc_command(srcs=SOURCES, out=OUT, extra=()):
    return [
        HIPCC,
        "--offload-arch=" + ARCH,
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"),
        *extra,
        *srcs,
        "-o", out,
    ]


def build(verbose=False):
    """Compile the library (always recompiles); raise on failure."""
    tmp = OUT + ".tmp"
    cmd = hipcc_command(out=tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building %s" % " ".join(SOURCES))
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
