"""Builds the HIP library fteik_amd/lib/libfteik_shared.so for gfx950.

    python -m fteik_amd.build          (or fteik_amd.build.build())

One hipcc link of all sources in fteik_amd/csrc.  The library keeps the
reference's file name so ctypes callers that search LD_LIBRARY_PATH for
libfteik_shared.so (python/fteik3d.py:26-37) find it unchanged.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libfteik_shared.so")
SOURCES = ["capi.cpp", "sweep3d.hip", "sweep2d.hip", "util.hip", "locate.hip",
           "stencil_eval.hip", "batch3d.hip", "surrogate.hip"]

# -ffp-contract=off: no FMA contraction, so the FP64 path keeps the
# reference's rounding; -packed-fp32-ops: no packed FP32 instructions.
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
         "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-Wno-unused-command-line-argument"]


def hipcc():
    for p in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if p and (os.path.isfile(p) or p == "hipcc"):
            return p
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES
            if os.path.isfile(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp%d" % os.getpid()
    cmd = [hipcc()] + FLAGS + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building %s" % LIB)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
