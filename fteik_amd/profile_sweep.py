"""Profiling driver for the 3D sweep kernels (run under rocprofv3).

    python3 -m fteik_amd.profile_sweep [--sweeps 16] [--copy-mib 1024]

Solves the 257^3 layered workload once in FP64, launches `--sweeps`
further Gauss-Seidel sweep kernels on the solved field (the kernels whose
HBM traffic and duration are measured), then copies `--copy-mib` MiB
device-to-device with hipMemcpy. The copy is the calibration: its traffic
is known (the buffer size is read once and written once), so the
counter-reported bytes of the copy give the scale factor for the sweep
kernels' counters.
"""
import argparse
import ctypes
import os
import sys
from ctypes import byref, c_double, c_int

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sweeps", type=int, default=16)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--precision", type=int, default=0)
    args = ap.parse_args()
    import fteik_amd
    from bench import layered_model, N, H
    L = fteik_amd.load_library()
    ierr = c_int(1)
    dp = ctypes.POINTER(c_double)
    vel, (zs, xs, ys) = layered_model()
    L.fteik_solver3d_initialize(N, N, N, 0.0, 0.0, 0.0, H, H, H, 2, 5.0, 0.0,
                                0, byref(ierr))
    assert ierr.value == 0
    L.fteik_solver3d_setVelocityModel64f(vel.size, vel.ctypes.data_as(dp),
                                         byref(ierr))
    z, x, y = (np.array([v]) for v in (zs, xs, ys))
    L.fteik_solver3d_setSources64f(1, z.ctypes.data_as(dp),
                                   x.ctypes.data_as(dp), y.ctypes.data_as(dp),
                                   byref(ierr))
    L.fteik_solver3d_setPrecision(args.precision, byref(ierr))
    L.fteik_solver3d_solveSourceLSM(1, byref(ierr))
    assert ierr.value == 0
    ms = c_double()
    nrep = max(1, args.sweeps // 8)
    L.fteik_amd_timeSweeps3d(args.precision, nrep, byref(ms), byref(ierr))
    assert ierr.value == 0
    print("sweep launches: %d, %.3f ms each" % (8 * nrep, ms.value),
          flush=True)
    hip = ctypes.CDLL("libamdhip64.so.7")
    n = args.copy_mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0   # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print("copy: %d MiB" % args.copy_mib, flush=True)


if __name__ == "__main__":
    main()
