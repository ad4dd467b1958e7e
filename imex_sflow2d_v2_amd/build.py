"""Build the gfx950 library ``libimexsw.so`` in-tree with hipcc.

FP policy: IEEE division and square root (hipcc defaults), no FMA
contraction (``-ffp-contract=off``) so device results follow the reference's
expression order, and no packed-FP32 instruction selection.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libimexsw.so")
SOURCES = ["imexsw.hip"]
HEADERS = ["imexsw_math.h", "imexsw_physics.h", "imexsw_newton.h", "imexsw_kernels.h",
           "imexsw_batch.h", "imexsw_mlp.h", "imexsw_q2d.h"]
ARCH = os.environ.get("IMEXSW_ARCH", "gfx950")

HIPCC_FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


def build(verbose=True, extra=()):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + list(extra) + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=CSRC)
    os.replace(tmp, LIB)
    return LIB


def stale():
    if not os.path.exists(LIB):
        return True
    t = os.path.getmtime(LIB)
    for f in SOURCES + HEADERS:
        if os.path.getmtime(os.path.join(CSRC, f)) > t:
            return True
    inc = os.path.join(os.path.dirname(HERE), "include", "imexsw.h")
    return os.path.getmtime(inc) > t


if __name__ == "__main__":
    build(extra=sys.argv[1:])
