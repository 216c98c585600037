"""fteik_amd: MI355X (gfx950) implementation of fteik's level-set fast
sweeping eikonal solvers (3D, 2D) and grid-search locator.

The product is the C-ABI library lib/libfteik_shared.so (declared in
include/fteik/fteik_amd.h), a drop-in for the reference's
libfteik_shared.so.  This module loads it with ctypes and mirrors the
reference's Python classes (python/fteik3d.py, python/fteik2d.py,
python/fteikAnalytic.py) with the same method names, argument meanings and
return conventions, plus a Locator class over the locate_* entry points.

There is no CPU path: if the library or a GPU is missing, loading or the
first call fails with an error.
"""
import ctypes
import os
from ctypes import POINTER, byref, c_bool, c_double, c_float, c_int

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libfteik_shared.so")

FTEIK_NATURAL_ORDERING = 0
FTEIK_ZXY_ORDERING = 0
FTEIK_ZX_ORDERING = 0
FTEIK_XYZ_ORDERING = 1
FTEIK_XZ_ORDERING = 1
FTEIK_ZYX_ORDERING = 2

_dp = POINTER(c_double)
_fp = POINTER(c_float)
_ip = POINTER(c_int)
_lib = None


def _d(a):
    return a.ctypes.data_as(_dp)


def _f(a):
    return a.ctypes.data_as(_fp)


def _i(a):
    return a.ctypes.data_as(_ip)


def library_path():
    return LIB_PATH


def load_library(path=None):
    """Loads libfteik_shared.so (built by fteik_amd.build).  Raises if it
    has not been built: the package has no fallback implementation."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.isfile(p):
        raise RuntimeError("fteik_amd: %s is missing; build it with "
                           "`python -m fteik_amd.build`" % p)
    L = ctypes.CDLL(p)
    i, d = c_int, c_double
    sig = {
        "fteik_solver3d_initialize": [i, i, i, d, d, d, d, d, d, i, d, d, i,
                                      _ip],
        "fteik_solver3d_setVelocityModel64f": [i, _dp, _ip],
        "fteik_solver3d_setNodalVelocityModel64f": [i, i, _dp, _ip],
        "fteik_solver3d_setCellVelocityModel64f": [i, i, _dp, _ip],
        "fteik_solver3d_setSources64f": [i, _dp, _dp, _dp, _ip],
        "fteik_solver3d_setReceivers64f": [i, _dp, _dp, _dp, _ip],
        "fteik_solver3d_solveSourceLSM": [i, _ip],
        "fteik_solver3d_getTravelTimes64f": [i, _dp, _ip],
        "fteik_solver3d_getTravelTimeField64f": [i, i, _dp, _ip],
        "fteik_solver3d_getTravelTimeFieldCell64f": [i, i, _dp, _ip],
        "fteik_solver3d_getNumberOfSources": [_ip, _ip],
        "fteik_solver3d_getNumberOfReceivers": [_ip, _ip],
        "fteik_solver3d_setNumberOfSweeps": [i, _ip],
        "fteik_solver3d_setSphereToCartEpsilon": [d, _ip],
        "fteik_solver3d_setVerbosity": [i],
        "fteik_solver3d_setConvergenceTolerance": [d],
        "fteik_solver3d_setPrecision": [i, _ip],
        "fteik_solver3d_setSchedule": [i, _ip],
        "fteik_solver3d_setMixedParameters": [d, i, _ip],
        "fteik_solver3d_setTile": [i, i, _ip],
        "fteik_solver3d_setTileSkipping": [i, _ip],
        "fteik_solver3d_setSweepOverlap": [i, _ip],
        "fteik_solver3d_setChainTile": [i, i, _ip],
        "fteik_solver3d_getSweepOverlap": [_ip, _ip, _ip, _ip],
        "fteik_solver3d_getTile": [_ip, _ip, _ip, _ip],
        "fteik_solver3d_autotune": [i, _ip, _ip, _ip],
        "fteik_batch3d_autotune": [i, d, _ip, _ip],
        "fteik_solver3d_getLastSolveInfo": [_dp, _ip, _dp, _ip],
        "fteik_solver3d_free": [],
        "fteik_amd_finalize": [],
        "fteik_solver3d_traceRaysToPoints": [i, _dp, _dp, _dp, _ip],
        "fteik_solver3d_traceRaysToReceivers": [_ip],
        "fteik_solver3d_getTravelTimeGradient64f": [i, _dp, _ip],
        "fteik_solver2d_traceRaysToPoints": [i, _dp, _dp, _ip],
        "fteik_solver2d_traceRaysToReceivers": [_ip],
        "fteik_solver2d_getTravelTimeGradient64f": [i, _dp, _dp, _ip],
        "fteik_rays_setODEIntegrator": [i, _ip],
        "fteik_rays_getNumberOfRays": [_ip, _ip],
        "fteik_rays_getTravelTimes64f": [i, _dp, _ip],
        "fteik_rays_getRayPath": [i, i, _dp, _ip, _dp, _ip],
        "fteik_model_index2gridF": [i, _ip, _ip, _ip, _ip],
        "fteik_solver2d_initialize": [i, i, d, d, d, d, i, d, d, i, _ip],
        "fteik_solver2d_setVelocityModel64f": [i, _dp, _ip],
        "fteik_solver2d_setNodalVelocityModel64f": [i, i, _dp, _ip],
        "fteik_solver2d_setCellVelocityModel64f": [i, i, _dp, _ip],
        "fteik_solver2d_setSources64f": [i, _dp, _dp, _ip],
        "fteik_solver2d_setReceivers64f": [i, _dp, _dp, _ip],
        "fteik_solver2d_solveSourceLSM": [i, _ip],
        "fteik_solver2d_solveSourceFSM": [i, _ip],
        "fteik_solver2d_solveLSM": [_ip],
        "fteik_solver2d_getTravelTimeField64f": [i, i, _dp, _ip],
        "fteik_solver2d_getTravelTimeFieldCell64f": [i, i, _dp, _ip],
        "fteik_solver2d_getTravelTimes64f": [i, _dp, _ip],
        "fteik_solver2d_getNumberOfReceivers": [_ip, _ip],
        "fteik_solver2d_free": [],
        "locate_initialize": [i, i, i, _ip],
        "locate_finalize": [],
        "locate_setTravelTimeField64f": [i, i, _dp, _ip],
        "locate_setTravelTimeField32f": [i, i, _fp, _ip],
        "locate_setObservation64f": [i, i, c_bool, d, _ip, _dp, _dp, _ip],
        "locate_locateEvent": [i, _ip, _dp, _dp],
        "locate_locateEvents": [i, _ip, _ip, _dp, _dp, _ip],
        "locate_locateEventsNoAlloc": [i, _ip, _ip, _dp, _dp, _ip],
        "locate_setReferenceCompatibility": [i, _ip],
        "fteik_analytic_initialize64f": [i, i, i, d, d, d, d, d, d, i, _ip],
        "fteik_analytic_setSources64f": [i, _dp, _dp, _dp, _ip],
        "fteik_analytic_setReceivers64f": [i, _dp, _dp, _dp, _ip],
        "fteik_analytic_setConstantVelocity64f": [d, _ip],
        "fteik_analytic_setLinearVelocityGradient64f": [d, d, _ip],
        "fteik_analytic_solveSourceConstantVelocity64f": [i, _ip],
        "fteik_analytic_solveSourceLinearVelocityGradient64f": [i, _ip],
        "fteik_analytic_getTravelTimeField64f": [i, i, _dp, _ip],
        "fteik_analytic_getTravelTimesConstantVel64f": [i, _dp, _ip],
        "fteik_analytic_getTravelTimesGradientVel64f": [i, _dp, _ip],
        "fteik_analytic_free": [],
        "fteik_graph3d_initialize": [i, i, i, i, _ip],
        "fteik_graph3d_makeLevelStructures": [_ip],
        "fteik_graph3d_getIJKVF": [i, i, _ip, _ip],
        "fteik_graph3d_getLevelPointerF": [i, _ip, _ip],
        "fteik_graph3d_getNumberOfLevels": [_ip],
        "fteik_graph3d_getMaxLevelSize": [_ip],
        "fteik_amd_evalLocalSolver3d": [i, _dp, _ip, _dp, _dp, _ip],
        "fteik_amd_prepareSource3d": [i, _ip],
        "fteik_amd_timeSweeps3d": [i, i, _dp, _ip],
        "fteik_amd_timeStencilChain3d": [i, i, _dp, _ip],
        "fteik_amd_sqrtStencil64f": [i, _dp, _dp, _ip],
        "fteik_solver2d_setSchedule": [i, _ip],
        "fteik_batch3d_initialize": [i, _ip],
        "fteik_batch3d_free": [],
        "fteik_batch3d_setGrid": [i, i, i, i, d, d, d, d, d, d, _dp, d, d, d,
                                  _ip],
        "fteik_batch3d_setPrecision": [i, _ip],
        "fteik_batch3d_setMixedParameters": [d, i, _ip],
        "fteik_batch3d_setTile": [i, _ip],
        "fteik_batch3d_solve": [i, d, _ip],
        "fteik_batch3d_getLastSolveInfo": [_dp, _ip, _dp, _ip],
        "fteik_batch3d_getTravelTimeField64f": [i, i, i, _dp, _ip],
        "fteik_batch3d_getNumberOfGrids": [_ip, _ip],
        "locate_setTravelTimeFieldFromBatch3d": [i, i, _ip],
        "locate_setTravelTimeFieldsFromStations3d": [i, _dp, _dp, _dp, i, d,
                                                     i, i, _ip],
        "fteik_surrogate_initialize": [i, i, i, i, d, d, d, _dp, _dp, _dp, d,
                                       d, i, _ip],
        "fteik_surrogate_free": [],
        "fteik_surrogate_getNumberOfParameters": [_ip, _ip],
        "fteik_surrogate_getParameters": [i, _fp, _ip],
        "fteik_surrogate_setParameters": [i, _fp, _ip],
        "fteik_surrogate_forward": [i, _fp, _fp, _ip],
        "fteik_surrogate_gradients": [i, _fp, _fp, _fp, _dp, _ip],
        "fteik_surrogate_train": [i, i, d, i, _dp, _ip],
        "fteik_surrogate_setNetworkProposals": [i, _ip],
        "locate_locateEventsSurrogate": [i, _ip, i, i, _ip, _dp, _dp, _ip,
                                         _ip],
        "fteik_amd_gemmBf16": [i, i, i, _fp, _fp, _fp, _ip],
        "fteik_amd_invHiddenBf16": [i, _fp, _fp, _fp,
                                    ctypes.POINTER(ctypes.c_ushort),
                                    ctypes.POINTER(ctypes.c_ushort), _ip],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = None
    L.fteik_graph3d_getNumberOfLevels.restype = c_int
    L.fteik_graph3d_getMaxLevelSize.restype = c_int
    L.fteik_amd_version.restype = ctypes.c_char_p
    if path is None:
        _lib = L
    return L


class FteikError(RuntimeError):
    pass


def _check(ierr, what):
    if ierr.value != 0:
        raise FteikError("%s failed (ierr=%d)" % (what, ierr.value))


def _as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ======================================================================
class fteik3d:
    """Mirror of the reference's python/fteik3d.py class over the MI355X
    library.  Methods return 0 on success and -1 on error, as there."""

    def __init__(self, fteik_library=None):
        self.fteik3d = load_library(fteik_library)
        self.linit = False
        self.nx = self.ny = self.nz = 0
        self.nsrc = self.nrec = 0

    def __enter__(self):
        return self

    def __exit__(self, *args):
        self.free()

    def free(self):
        self.nx = self.ny = self.nz = self.nsrc = self.nrec = 0
        self.linit = False
        self.fteik3d.fteik_solver3d_free()

    def initialize(self, nx, ny, nz, dx, dy, dz, x0=0.0, y0=0.0, z0=0.0,
                   nsweep=2, eps=5.0, convTol=0.0, verbose=0):
        ierr = c_int(1)
        self.fteik3d.fteik_solver3d_initialize(nz, nx, ny, z0, x0, y0, dz, dx,
                                               dy, nsweep, eps, convTol,
                                               verbose, byref(ierr))
        if ierr.value != 0:
            print("Error initializing solver")
            return -1
        self.nx, self.ny, self.nz = nx, ny, nz
        self.linit = True
        return 0

    def setVelocityModel(self, vel):
        """[nz-1 x ny-1 x nx-1] cell or [nz x ny x nx] nodal model, m/s
        (FTEIK_ZYX_ORDERING after a Fortran-order flatten, as the
        reference's binding does)."""
        vel = np.asarray(vel)
        v = _as_f64(np.reshape(vel, vel.size, order="F"))
        ncell = (self.nx - 1) * (self.ny - 1) * (self.nz - 1)
        ierr = c_int(1)
        if v.size != ncell and v.size == self.nx * self.ny * self.nz:
            self.fteik3d.fteik_solver3d_setNodalVelocityModel64f(
                v.size, FTEIK_ZYX_ORDERING, _d(v), byref(ierr))
        else:
            self.fteik3d.fteik_solver3d_setCellVelocityModel64f(
                v.size, FTEIK_ZYX_ORDERING, _d(v), byref(ierr))
        if ierr.value != 0:
            print("Error setting velocity")
            return -1
        return 0

    def setSources(self, xsrc, ysrc, zsrc):
        x, y, z = _as_f64(np.atleast_1d(xsrc)), _as_f64(np.atleast_1d(ysrc)), \
            _as_f64(np.atleast_1d(zsrc))
        nsrc = min(len(x), len(y), len(z))
        ierr = c_int(1)
        self.fteik3d.fteik_solver3d_setSources64f(nsrc, _d(z), _d(x), _d(y),
                                                  byref(ierr))
        if ierr.value != 0:
            print("Error setting sources")
            return -1
        self.nsrc = nsrc
        return 0

    def setReceivers(self, xrec, yrec, zrec):
        x, y, z = _as_f64(np.atleast_1d(xrec)), _as_f64(np.atleast_1d(yrec)), \
            _as_f64(np.atleast_1d(zrec))
        nrec = min(len(x), len(y), len(z))
        ierr = c_int(1)
        self.fteik3d.fteik_solver3d_setReceivers64f(nrec, _d(z), _d(x), _d(y),
                                                    byref(ierr))
        if ierr.value != 0:
            print("Error setting receivers")
            return -1
        self.nrec = nrec
        return 0

    def solveSource(self, isrc):
        """Solves source isrc (1-based); returns 0 or -1."""
        ierr = c_int(1)
        self.fteik3d.fteik_solver3d_solveSourceLSM(isrc, byref(ierr))
        return 0 if ierr.value == 0 else -1

    def solveLSM(self):
        errorAll = 0
        for i in range(self.nsrc):
            if self.solveSource(i + 1) != 0:
                print("Failed to solve for source %d" % (i + 1))
                errorAll += 1
        return errorAll

    def getTravelTimeField(self, wantCell=False, order=FTEIK_ZYX_ORDERING):
        if wantCell:
            n = (self.nz - 1) * (self.nx - 1) * (self.ny - 1)
        else:
            n = self.nz * self.nx * self.ny
        tt = np.zeros(n)
        ierr = c_int(1)
        if wantCell:
            self.fteik3d.fteik_solver3d_getTravelTimeFieldCell64f(
                n, order, _d(tt), byref(ierr))
        else:
            self.fteik3d.fteik_solver3d_getTravelTimeField64f(
                n, order, _d(tt), byref(ierr))
        if ierr.value != 0:
            print("Error getting travel time field")
            return None
        if order != FTEIK_ZYX_ORDERING:
            return tt
        shape = [self.nz - 1, self.ny - 1, self.nx - 1] if wantCell else \
            [self.nz, self.ny, self.nx]
        return np.reshape(tt, shape, order="F")

    def getTravelTimes(self):
        if self.nrec < 1:
            print("No receiver locations set")
            return None
        ttr = np.zeros(self.nrec)
        ierr = c_int(1)
        self.fteik3d.fteik_solver3d_getTravelTimes64f(self.nrec, _d(ttr),
                                                      byref(ierr))
        if ierr.value != 0:
            print("Error getting travel times")
            return None
        return ttr

    # -- extensions -------------------------------------------------------
    def setPrecision(self, mode):
        ierr = c_int(1)
        self.fteik3d.fteik_solver3d_setPrecision(mode, byref(ierr))
        return 0 if ierr.value == 0 else -1

    def setSchedule(self, schedule):
        ierr = c_int(1)
        self.fteik3d.fteik_solver3d_setSchedule(schedule, byref(ierr))
        return 0 if ierr.value == 0 else -1

    def lastSolveInfo(self):
        ms, upd = c_double(), c_double()
        corr, ierr = c_int(), c_int()
        self.fteik3d.fteik_solver3d_getLastSolveInfo(byref(ms), byref(corr),
                                                     byref(upd), byref(ierr))
        return {"ms": ms.value, "corr_iterations": corr.value,
                "updates": upd.value}


# ======================================================================
class fteik2d:
    """Mirror of the reference's python/fteik2d.py class."""

    def __init__(self, fteik_library=None):
        self.fteik2d = load_library(fteik_library)
        self.nx = self.nz = self.nsrc = self.nrec = 0

    def __enter__(self):
        return self

    def __exit__(self, *args):
        self.free()

    def free(self):
        self.nx = self.nz = self.nsrc = self.nrec = 0
        self.fteik2d.fteik_solver2d_free()

    def finalize(self):
        self.free()

    def initialize(self, nx, nz, dx, dz, x0=0.0, z0=0.0, nsweep=2, eps=5.0,
                   convTol=0.0, verbose=0):
        ierr = c_int(1)
        self.fteik2d.fteik_solver2d_initialize(nz, nx, z0, x0, dz, dx, nsweep,
                                               eps, convTol, verbose,
                                               byref(ierr))
        if ierr.value != 0:
            print("Error initializing solver")
            return -1
        self.nx, self.nz, self.nsrc = nx, nz, 0
        return 0

    def setVelocityModel(self, vel):
        vel = np.asarray(vel)
        v = _as_f64(np.reshape(vel, vel.size, order="F"))
        ierr = c_int(1)
        if v.size != (self.nx - 1) * (self.nz - 1) and \
                v.size == self.nx * self.nz:
            self.fteik2d.fteik_solver2d_setNodalVelocityModel64f(
                v.size, FTEIK_ZX_ORDERING, _d(v), byref(ierr))
        else:
            self.fteik2d.fteik_solver2d_setCellVelocityModel64f(
                v.size, FTEIK_ZX_ORDERING, _d(v), byref(ierr))
        if ierr.value != 0:
            print("Error setting velocity")
            return -1
        return 0

    def solveLSM(self):
        ierr = c_int(1)
        self.fteik2d.fteik_solver2d_solveLSM(byref(ierr))
        if ierr.value != 0:
            print("Error solving eikonal equation")
            return -1
        return 0

    def solveSource(self, isrc, fsm=False):
        ierr = c_int(1)
        fn = self.fteik2d.fteik_solver2d_solveSourceFSM if fsm else \
            self.fteik2d.fteik_solver2d_solveSourceLSM
        fn(isrc, byref(ierr))
        return 0 if ierr.value == 0 else -1

    def getTravelTimeField(self, wantCell=False):
        n = (self.nx - 1) * (self.nz - 1) if wantCell else self.nx * self.nz
        tt = np.zeros(n)
        ierr = c_int(1)
        if wantCell:
            self.fteik2d.fteik_solver2d_getTravelTimeFieldCell64f(
                n, FTEIK_ZX_ORDERING, _d(tt), byref(ierr))
        else:
            self.fteik2d.fteik_solver2d_getTravelTimeField64f(
                n, FTEIK_ZX_ORDERING, _d(tt), byref(ierr))
        if ierr.value != 0:
            print("Error getting travel time field")
            return None
        shape = [self.nz - 1, self.nx - 1] if wantCell else [self.nz, self.nx]
        return np.reshape(tt, shape, order="F")

    def getTravelTimes(self):
        if self.nrec < 1:
            print("No receiver locations set")
            return None
        ttr = np.zeros(self.nrec * max(self.nsrc, 1))
        ierr = c_int(1)
        self.fteik2d.fteik_solver2d_getTravelTimes64f(self.nrec, _d(ttr),
                                                      byref(ierr))
        if ierr.value != 0:
            print("Error getting travel times")
            return None
        return np.reshape(ttr, [self.nrec, self.nsrc], order="F")

    def setSources(self, xsrc, zsrc):
        x, z = _as_f64(np.atleast_1d(xsrc)), _as_f64(np.atleast_1d(zsrc))
        nsrc = min(len(x), len(z))
        ierr = c_int(1)
        self.fteik2d.fteik_solver2d_setSources64f(nsrc, _d(z), _d(x),
                                                  byref(ierr))
        if ierr.value != 0:
            print("Error setting sources")
            return -1
        self.nsrc = nsrc
        return 0

    def setReceivers(self, xrec, zrec):
        x, z = _as_f64(np.atleast_1d(xrec)), _as_f64(np.atleast_1d(zrec))
        nrec = min(len(x), len(z))
        ierr = c_int(1)
        self.fteik2d.fteik_solver2d_setReceivers64f(nrec, _d(z), _d(x),
                                                    byref(ierr))
        if ierr.value != 0:
            print("Error setting receivers")
            return -1
        self.nrec = nrec
        return 0


# ======================================================================
class fteikAnalytic:
    """Mirror of the reference's python/fteikAnalytic.py class."""

    def __init__(self, fteik_library=None):
        self.lib = load_library(fteik_library)
        self.nx = self.ny = self.nz = self.nsrc = self.nrec = 0

    def free(self):
        self.lib.fteik_analytic_free()

    def initialize(self, nx, ny, nz, dx, dy, dz, x0=0.0, y0=0.0, z0=0.0,
                   verbose=0):
        ierr = c_int(1)
        self.lib.fteik_analytic_initialize64f(nz, nx, ny, z0, x0, y0, dz, dx,
                                              dy, verbose, byref(ierr))
        if ierr.value != 0:
            return -1
        self.nx, self.ny, self.nz = nx, ny, nz
        return 0

    def setSources(self, xsrc, ysrc, zsrc):
        x, y, z = (_as_f64(np.atleast_1d(v)) for v in (xsrc, ysrc, zsrc))
        ierr = c_int(1)
        self.lib.fteik_analytic_setSources64f(len(x), _d(z), _d(x), _d(y),
                                              byref(ierr))
        if ierr.value != 0:
            return -1
        self.nsrc = len(x)
        return 0

    def setReceivers(self, xrec, yrec, zrec):
        x, y, z = (_as_f64(np.atleast_1d(v)) for v in (xrec, yrec, zrec))
        ierr = c_int(1)
        self.lib.fteik_analytic_setReceivers64f(len(x), _d(z), _d(x), _d(y),
                                                byref(ierr))
        if ierr.value != 0:
            return -1
        self.nrec = len(x)
        return 0

    def setConstantVelocity(self, vel=5000.0):
        ierr = c_int(1)
        self.lib.fteik_analytic_setConstantVelocity64f(vel, byref(ierr))
        return 0 if ierr.value == 0 else -1

    def setLinearVelocityGradient(self, v1=5000.0, v2=6000.0):
        ierr = c_int(1)
        self.lib.fteik_analytic_setLinearVelocityGradient64f(v1, v2,
                                                             byref(ierr))
        return 0 if ierr.value == 0 else -1

    def solveConstantVelocity(self):
        ierr = c_int(1)
        for i in range(self.nsrc):
            self.lib.fteik_analytic_solveSourceConstantVelocity64f(
                i + 1, byref(ierr))
        return 0 if ierr.value == 0 else -1

    def solveLinearVelocityGradient(self):
        ierr = c_int(1)
        for i in range(self.nsrc):
            self.lib.fteik_analytic_solveSourceLinearVelocityGradient64f(
                i + 1, byref(ierr))
        return 0 if ierr.value == 0 else -1

    def getTravelTimeField(self, order=FTEIK_ZYX_ORDERING):
        n = self.nz * self.nx * self.ny
        tt = np.zeros(n)
        ierr = c_int(1)
        self.lib.fteik_analytic_getTravelTimeField64f(n, order, _d(tt),
                                                      byref(ierr))
        if ierr.value != 0:
            return None
        if order != FTEIK_ZYX_ORDERING:
            return tt
        if self.ny > 1:
            return np.reshape(tt, [self.nz, self.ny, self.nx], order="F")
        return np.reshape(tt, [self.nz, self.nx], order="F")

    def _rec(self, fn):
        ttr = np.zeros(max(self.nrec, 1) * max(self.nsrc, 1))
        ierr = c_int(1)
        fn(self.nrec, _d(ttr), byref(ierr))
        if ierr.value != 0:
            return None
        return np.reshape(ttr, [self.nrec, self.nsrc], order="F")

    def getTravelTimesConstantVelocity(self):
        return self._rec(self.lib.fteik_analytic_getTravelTimesConstantVel64f)

    def getTravelTimesGradientVelocity(self):
        return self._rec(self.lib.fteik_analytic_getTravelTimesGradientVel64f)


# ======================================================================
class Locator:
    """locate_* (locate.F90) over the MI355X library."""

    def __init__(self, nEvents, ntf, ngrd, fteik_library=None):
        self.lib = load_library(fteik_library)
        ierr = c_int(1)
        self.lib.locate_initialize(nEvents, ntf, ngrd, byref(ierr))
        _check(ierr, "locate_initialize")
        self.nEvents, self.ntf, self.ngrd = nEvents, ntf, ngrd

    def finalize(self):
        self.lib.locate_finalize()

    def setReferenceCompatibility(self, mode):
        ierr = c_int(1)
        self.lib.locate_setReferenceCompatibility(int(mode), byref(ierr))
        _check(ierr, "locate_setReferenceCompatibility")

    def setTravelTimeField(self, itf, tt):
        tt = np.asarray(tt)
        ierr = c_int(1)
        if tt.dtype == np.float32:
            t = np.ascontiguousarray(tt)
            self.lib.locate_setTravelTimeField32f(t.size, itf, _f(t),
                                                  byref(ierr))
        else:
            t = _as_f64(tt)
            self.lib.locate_setTravelTimeField64f(t.size, itf, _d(t),
                                                  byref(ierr))
        _check(ierr, "locate_setTravelTimeField")

    def setObservation(self, evnmbr, obs2tf, picks, wts=None, t0=None):
        o2t = np.ascontiguousarray(obs2tf, dtype=np.int32)
        p = _as_f64(picks)
        w = _as_f64(np.ones_like(p) if wts is None else wts)
        ierr = c_int(1)
        self.lib.locate_setObservation64f(evnmbr, p.size, t0 is not None,
                                          0.0 if t0 is None else t0, _i(o2t),
                                          _d(p), _d(w), byref(ierr))
        _check(ierr, "locate_setObservation64f")

    def locateEvent(self, evnmbr):
        opt, t0, obj = c_int(), c_double(), c_double()
        self.lib.locate_locateEvent(evnmbr, byref(opt), byref(t0), byref(obj))
        return opt.value, t0.value, obj.value

    def locateEvents(self, evnmbrs):
        ev = np.ascontiguousarray(evnmbrs, dtype=np.int32)
        n = ev.size
        opt = np.zeros(n, dtype=np.int32)
        t0 = np.zeros(n)
        obj = np.zeros(n)
        ierr = c_int(1)
        self.lib.locate_locateEvents(n, _i(ev), _i(opt), _d(t0), _d(obj),
                                     byref(ierr))
        _check(ierr, "locate_locateEvents")
        return opt, t0, obj
