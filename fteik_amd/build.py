"""Builds the HIP library fteik_amd/lib/libfteik_shared.so for gfx950.

    python -m fteik_amd.build          (or fteik_amd.build.build())

Every source in fteik_amd/csrc is compiled to an object of its own, in
parallel, and the objects are linked with hipcc.  The library keeps the
reference's file name so ctypes callers that search LD_LIBRARY_PATH for
libfteik_shared.so (python/fteik3d.py:26-37) find it unchanged.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(LIBDIR, "obj")
LIB = os.path.join(LIBDIR, "libfteik_shared.so")
SOURCES = ["capi_common.cpp", "capi3d.cpp", "capi2d.cpp", "capi_locate.cpp",
           "capi_model.cpp", "capi_batch.cpp", "capi_surrogate.cpp",
           "capi_rays.cpp", "locate_noalloc.hip",
           "sweep3d.hip", "sweep2d.hip", "util.hip", "locate.hip",
           "stencil_eval.hip", "batch3d.hip", "surrogate.hip", "rays.hip"]

# -ffp-contract=off: no FMA contraction, so the FP64 path keeps the
# reference's rounding; -packed-fp32-ops: no packed FP32 instructions.
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
         "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-Wno-unused-command-line-argument"]


def hipcc():
    for p in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if p and (os.path.isfile(p) or p == "hipcc"):
            return p
    raise RuntimeError("hipcc not found")


def _compile(src, verbose):
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    tmp = obj + ".tmp%d" % os.getpid()
    cmd = [hipcc()] + FLAGS + ["-c", "-o", tmp, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return None, r.stdout + r.stderr
    os.replace(tmp, obj)
    return obj, ""


def build(verbose=False, jobs=None):
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES
            if os.path.isfile(os.path.join(CSRC, s))]
    jobs = jobs or min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    objs, errs = [], []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, err in ex.map(lambda s: _compile(s, verbose), srcs):
            if obj is None:
                errs.append(err)
            else:
                objs.append(obj)
    if errs:
        sys.stderr.write("\n".join(errs))
        raise RuntimeError("hipcc failed building %s" % LIB)
    tmp = LIB + ".tmp%d" % os.getpid()
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed linking %s" % LIB)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
