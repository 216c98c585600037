"""Helpers shared by the GPU parity tests: drive the MI355X library through
its C-ABI (ctypes) and the CPU oracle on the same inputs."""
import ctypes
from ctypes import byref, c_double, c_int

import numpy as np

import fteik_amd as fa
from oracle import oracle as orc


def lib():
    return fa.load_library()


def _d(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _i(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def grid9(nz, nx, ny, h, z0=0.0, x0=0.0, y0=0.0):
    return (nz, nx, ny, h, h, h, z0, x0, y0)


def hetero(nz, nx, ny, h, seed):
    from tests.golden.gen_golden_models import hetero_model3d
    return hetero_model3d(np.random.default_rng(seed), nz, nx, ny, h)


class GPU3D:
    """One configured 3D solver in the library (process-global state)."""

    def __init__(self, nz, nx, ny, h, nsweep=2, eps=5.0, conv=0.0,
                 origin=(0.0, 0.0, 0.0), schedule=0, precision=0):
        self.L = lib()
        self.nz, self.nx, self.ny, self.h = nz, nx, ny, h
        ierr = c_int(1)
        z0, x0, y0 = origin
        self.L.fteik_solver3d_initialize(nz, nx, ny, z0, x0, y0, h, h, h,
                                         nsweep, eps, conv, 0, byref(ierr))
        assert ierr.value == 0
        self.L.fteik_solver3d_setSchedule(schedule, byref(ierr))
        assert ierr.value == 0
        self.L.fteik_solver3d_setPrecision(precision, byref(ierr))
        assert ierr.value == 0

    def set_cells(self, vel):
        v = np.ascontiguousarray(vel, dtype=np.float64)
        ierr = c_int(1)
        self.L.fteik_solver3d_setVelocityModel64f(v.size, _d(v), byref(ierr))
        assert ierr.value == 0

    def set_sources(self, zs, xs, ys):
        z, x, y = (np.ascontiguousarray(np.atleast_1d(a), dtype=np.float64)
                   for a in (zs, xs, ys))
        ierr = c_int(1)
        self.L.fteik_solver3d_setSources64f(z.size, _d(z), _d(x), _d(y),
                                            byref(ierr))
        assert ierr.value == 0

    def set_receivers(self, zr, xr, yr):
        z, x, y = (np.ascontiguousarray(a, dtype=np.float64)
                   for a in (zr, xr, yr))
        ierr = c_int(1)
        self.L.fteik_solver3d_setReceivers64f(z.size, _d(z), _d(x), _d(y),
                                              byref(ierr))
        assert ierr.value == 0
        self.nrec = z.size

    def solve(self, isrc=1):
        ierr = c_int(1)
        self.L.fteik_solver3d_solveSourceLSM(isrc, byref(ierr))
        return ierr.value

    def field(self, order=0):
        n = self.nz * self.nx * self.ny
        tt = np.zeros(n)
        ierr = c_int(1)
        self.L.fteik_solver3d_getTravelTimeField64f(n, order, _d(tt),
                                                    byref(ierr))
        assert ierr.value == 0
        return tt

    def cells(self, order=0):
        n = (self.nz - 1) * (self.nx - 1) * (self.ny - 1)
        tt = np.zeros(n)
        ierr = c_int(1)
        self.L.fteik_solver3d_getTravelTimeFieldCell64f(n, order, _d(tt),
                                                        byref(ierr))
        assert ierr.value == 0
        return tt

    def receivers(self):
        ttr = np.zeros(self.nrec)
        ierr = c_int(1)
        self.L.fteik_solver3d_getTravelTimes64f(self.nrec, _d(ttr),
                                                byref(ierr))
        assert ierr.value == 0
        return ttr

    def info(self):
        ms, upd = c_double(), c_double()
        corr, ierr = c_int(), c_int()
        self.L.fteik_solver3d_getLastSolveInfo(byref(ms), byref(corr),
                                               byref(upd), byref(ierr))
        return ms.value, corr.value, upd.value


def oracle3d(vel, nz, nx, ny, h, zs, xs, ys, nsweep, eps, origin=(0, 0, 0)):
    """Oracle field (natural order) for the same inputs; schedule 1 is the
    OpenMP level-synchronous loop, bitwise equal to the serial one."""
    slow = orc.cell_slowness(vel, nz, nx, ny)
    g = grid9(nz, nx, ny, h, *origin)
    src = orc.source_locate(g, zs, xs, ys, slow)
    return orc.solve3d(slow, nz, nx, ny, h, h, h, src, nsweep, eps,
                       schedule=1), slow, src
