"""Build the gfx950 library ``libimexsw.so`` in-tree with hipcc.

The library is one C-ABI translation unit (``imexsw.hip``) plus
``imexsw_ns.hip`` compiled once per number of solid classes (0..6), which
holds the kernels templated on it; the objects compile in parallel and are
linked into one shared library.

FP policy: IEEE division and square root (hipcc defaults), no FMA
contraction (``-ffp-contract=off``) so device results follow the reference's
expression order, and no packed-FP32 instruction selection.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libimexsw.so")
NS_VALUES = range(0, 7)
SOURCES = ["imexsw.hip", "imexsw_ns.hip"]
HEADERS = ["imexsw_math.h", "imexsw_physics.h", "imexsw_newton.h", "imexsw_kernels.h",
           "imexsw_batch.h", "imexsw_mlp.h", "imexsw_q2d.h", "imexsw_ctx.h",
           "imexsw_stages.h", "imexsw_profiles.h"]
ARCH = os.environ.get("IMEXSW_ARCH", "gfx950")

HIPCC_FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
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


def _jobs():
    n = os.cpu_count() or 4
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    return max(1, min(n, 16, int(os.environ.get("MAX_JOBS", "16"))))


def compile_units(extra=()):
    """(command, object) for every translation unit."""
    units = [([hipcc()] + HIPCC_FLAGS + list(extra) + ["-c", os.path.join(CSRC, "imexsw.hip"),
                                                      "-o", os.path.join(OBJ, "imexsw.o")],
              os.path.join(OBJ, "imexsw.o"))]
    for ns in NS_VALUES:
        o = os.path.join(OBJ, "imexsw_ns%d.o" % ns)
        units.append(([hipcc()] + HIPCC_FLAGS + list(extra) +
                      ["-DIMEXSW_NS=%d" % ns, "-c", os.path.join(CSRC, "imexsw_ns.hip"), "-o", o], o))
    return units


def build(verbose=True, extra=()):
    os.makedirs(OBJ, exist_ok=True)
    units = compile_units(extra)

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd, cwd=CSRC)

    with ThreadPoolExecutor(max_workers=_jobs()) as pool:
        for f in [pool.submit(run, cmd) for cmd, _ in units]:
            f.result()
    tmp = LIB + ".tmp"
    link = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + [o for _, o in units]
    run(link)
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
