"""ctypes front end to the CPU oracle (liboracle.so) and to the reference
build (oracle/_ref/libfteikref.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; the product package fteik_amd
never imports it.

Array conventions follow the reference: travel-time fields are z-fastest
("natural" order, index (iy*nx + ix)*nz + iz), slowness is z-fastest over
cells, grid indices exchanged with the C functions are 1-based.
"""
import ctypes
import os
from ctypes import POINTER, byref, c_double, c_float, c_int, c_ubyte

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libfteikref.so")
HUGE = 99999.0

_dp = POINTER(c_double)
_fp = POINTER(c_float)
_ip = POINTER(c_int)
_bp = POINTER(c_ubyte)


def _d(a):
    return a.ctypes.data_as(_dp)


def _f(a):
    return a.ctypes.data_as(_fp)


def _i(a):
    return a.ctypes.data_as(_ip)


class Source(ctypes.Structure):
    _fields_ = [("zsa", c_double), ("xsa", c_double), ("ysa", c_double),
                ("zsi", c_int), ("xsi", c_int), ("ysi", c_int),
                ("szero", c_double)]


class Consts3d(ctypes.Structure):
    _fields_ = [(n, c_double) for n in (
        "dz dx dy dzi dxi dyi dz2 dx2 dy2 dz2i dx2i dy2i dsum dsum_nine dsumi "
        "epsSolver dz2dx2 dz2dy2 dx2dy2 dz2i_dx2i dz2i_dy2i dx2i_dy2i "
        "dz2i_p_dx2i dz2i_p_dy2i dx2i_p_dy2i dz2i_p_dy2i_inv dz2i_p_dx2i_inv "
        "dx2i_p_dy2i_inv szero szero2 zsa xsa ysa").split()] + \
        [("zsi", c_int), ("xsi", c_int), ("ysi", c_int)]


_lib = None


def lib():
    """Load (building if needed) liboracle.so."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.isfile(ORACLE_SO):
        raise RuntimeError("oracle library missing: run make -C oracle")
    L = ctypes.CDLL(ORACLE_SO)
    L.or_source_locate.argtypes = [c_int, c_int, c_int, c_int, c_double,
                                   c_double, c_double, c_double, c_double,
                                   c_double, c_double, c_double, c_double,
                                   _dp, POINTER(Source)]
    L.or_consts3d_init.argtypes = [c_double, c_double, c_double, c_double,
                                   POINTER(Source), POINTER(Consts3d)]
    L.or_noinit3d.argtypes = [POINTER(Consts3d), _dp, _dp, c_double]
    L.or_noinit3d.restype = c_double
    L.or_init3d.argtypes = [POINTER(Consts3d), _dp, _dp, c_double, c_int,
                            c_int, c_int, c_int]
    L.or_init3d.restype = c_double
    L.or_tana3d.argtypes = [c_int, c_int, c_int, POINTER(Consts3d)]
    L.or_tana3d.restype = c_double
    L.or_seeds3d.argtypes = [POINTER(Consts3d), c_int, c_int, _ip, _dp]
    L.or_sweep_limits.argtypes = [c_int] * 8 + [_ip]
    L.or_levels3d.argtypes = [c_int, c_int, c_int, c_int, _ip, _ip]
    L.or_gather_slow3d.argtypes = [c_int] * 7 + [_dp, _dp]
    L.or_gather_tt3d.argtypes = [c_int] * 6 + [_dp, _dp, _dp]
    L.or_solve3d.argtypes = [c_int, c_int, c_int, c_double, c_double,
                             c_double, _dp, POINTER(Source), c_int, c_double,
                             c_int, _dp]
    L.or_sweep3d.argtypes = [c_int] * 6 + [_dp, POINTER(Consts3d), _dp]
    L.or_field_reorder3d.argtypes = [c_int] * 4 + [_dp, _dp]
    L.or_field_cell3d.argtypes = [c_int] * 4 + [_dp, _dp]
    L.or_cell_slowness.argtypes = [c_int] * 6 + [_dp, _dp]
    L.or_nodal_slowness.argtypes = [c_int] * 6 + [_dp, _dp]
    L.or_velocity_slowness.argtypes = [c_int, _dp, _dp]
    L.or_receivers_locate.argtypes = [c_int] * 4 + [c_double] * 6 + \
        [c_int, _dp, _dp, _dp, _ip, _dp]
    L.or_receivers_interp.argtypes = [c_int, c_int, c_int, c_double,
                                      c_double, c_double, c_int, _ip, _dp,
                                      _dp, _dp]
    L.or_solve2d.argtypes = [c_int, c_int, c_double, c_double, _dp,
                             POINTER(Source), c_int, c_double, _dp]
    L.or_noinit2d.argtypes = [_dp, _dp, c_double, c_double]
    L.or_noinit2d.restype = c_double
    L.or_init2d.argtypes = [c_int] * 5 + [c_double, c_double, c_double,
                                          POINTER(Source), _dp, _dp]
    L.or_init2d.restype = c_double
    L.or_analytic_constant.argtypes = [c_int] * 3 + [c_double] * 10 + [_dp]
    L.or_analytic_gradient.argtypes = [c_int] * 3 + [c_double] * 11 + [_dp]
    L.or_padlength32.argtypes = [c_int, c_int]
    L.or_locate_observation.argtypes = [c_int, _ip, _dp, _dp, c_int, _dp,
                                        _fp, _fp, _bp]
    L.or_locate_event.argtypes = [c_int, c_int, c_int, _fp, _fp, _fp, _bp,
                                  c_int, c_double, c_int, c_int, _fp, _fp,
                                  c_int, _dp, _dp]
    L.or_locate_event.restype = c_int
    _lib = L
    return L


# ------------------------------------------------------------------ model
def cell_slowness(vel, nz, nx, ny=None, order=0):
    """model.F90:469-559 -> slowness, natural cell order."""
    is3d = ny is not None
    ncell = (nz - 1) * (nx - 1) * ((ny - 1) if is3d else 1)
    v = np.ascontiguousarray(vel, dtype=np.float64).ravel(order="F") \
        if np.ndim(vel) > 1 else np.ascontiguousarray(vel, dtype=np.float64)
    s = np.zeros(ncell)
    ierr = lib().or_cell_slowness(int(is3d), nz, nx, ny or 1, order, v.size,
                                  _d(v), _d(s))
    if ierr:
        raise ValueError("invalid velocity model")
    return s


def nodal_slowness(vel, nz, nx, ny=None, order=0):
    """model.F90:311-440 (reference quirks kept)."""
    is3d = ny is not None
    ncell = (nz - 1) * (nx - 1) * ((ny - 1) if is3d else 1)
    v = np.ascontiguousarray(vel, dtype=np.float64).ravel()
    s = np.zeros(ncell)
    ierr = lib().or_nodal_slowness(int(is3d), nz, nx, ny or 1, order, v.size,
                                   _d(v), _d(s))
    if ierr:
        raise ValueError("invalid nodal velocity model")
    return s


def source_locate(grid, zs, xs, ys=0.0, slow=None, is3d=True):
    nz, nx, ny, dz, dx, dy, z0, x0, y0 = grid
    src = Source()
    sp = _d(slow) if slow is not None else None
    ierr = lib().or_source_locate(int(is3d), nz, nx, ny, dz, dx, dy, z0, x0,
                                  y0, zs, xs, ys, sp, byref(src))
    if ierr:
        raise ValueError("source out of bounds")
    return src


def consts3d(dz, dx, dy, eps, src):
    c = Consts3d()
    lib().or_consts3d_init(dz, dx, dy, eps, byref(src), byref(c))
    return c


# --------------------------------------------------------------- 3D solve
def solve3d(slow, nz, nx, ny, dz, dx, dy, src, nsweep, eps, schedule=0,
            nthreads=None):
    """solver3D.F90:1156-1243.  schedule 0: lexicographic (serial);
    schedule 1: level-synchronous with OpenMP (the reference's schedule)."""
    if nthreads is not None:
        os.environ["OMP_NUM_THREADS"] = str(nthreads)
    slow = np.ascontiguousarray(slow, dtype=np.float64)
    tt = np.empty(nz * nx * ny)
    ierr = lib().or_solve3d(nz, nx, ny, dz, dx, dy, _d(slow), byref(src),
                            nsweep, eps, schedule, _d(tt))
    if ierr:
        raise RuntimeError("oracle 3D solve failed")
    return tt


def sweep_limits(sweep, linit, nz, nx, ny, zsi, xsi, ysi):
    lim = np.zeros(6, dtype=np.int32)
    lib().or_sweep_limits(sweep, int(linit), nz, nx, ny, zsi, xsi, ysi,
                          _i(lim))
    return lim


def levels3d(sweep, nz, nx, ny):
    nlev = nx + ny + nz - 2
    ptr = np.zeros(nlev + 1, dtype=np.int32)
    ijkv = np.zeros(4 * nz * nx * ny, dtype=np.int32)
    if lib().or_levels3d(sweep, nz, nx, ny, _i(ptr), _i(ijkv)):
        raise ValueError("bad grid")
    return ptr, ijkv


def gather_slow3d(sweep, i, j, k, nz, nx, ny, slow):
    s7 = np.zeros(7)
    lib().or_gather_slow3d(sweep, i, j, k, nz, nx, ny, _d(slow), _d(s7))
    return s7


def gather_tt3d(sweep, i, j, k, nz, nx, tt):
    t7 = np.zeros(7)
    t8 = np.zeros(1)
    lib().or_gather_tt3d(sweep, i, j, k, nz, nx, _d(tt), _d(t7), _d(t8))
    return t7, t8[0]


def field_reorder3d(tt, nz, nx, ny, order):
    out = np.empty(nz * nx * ny)
    lib().or_field_reorder3d(nz, nx, ny, order, _d(tt), _d(out))
    return out


def field_cell3d(tt, nz, nx, ny, order):
    out = np.empty((nz - 1) * (nx - 1) * (ny - 1))
    lib().or_field_cell3d(nz, nx, ny, order, _d(tt), _d(out))
    return out


# ------------------------------------------------------------- receivers
def receivers(tt, grid, zr, xr, yr=None, is3d=True):
    nz, nx, ny, dz, dx, dy, z0, x0, y0 = grid
    zr = np.ascontiguousarray(zr, dtype=np.float64)
    xr = np.ascontiguousarray(xr, dtype=np.float64)
    yr = np.ascontiguousarray(yr if yr is not None else np.zeros_like(zr),
                              dtype=np.float64)
    n = zr.size
    ri = np.zeros(3 * n, dtype=np.int32)
    rd = np.zeros(3 * n)
    lib().or_receivers_locate(int(is3d), nz, nx, ny, dz, dx, dy, z0, x0, y0,
                              n, _d(zr), _d(xr), _d(yr), _i(ri), _d(rd))
    ttr = np.zeros(n)
    lib().or_receivers_interp(int(is3d), nz, nx, dz, dx, dy if is3d else 1.0,
                              n, _i(ri), _d(rd), _d(tt), _d(ttr))
    return ttr


# ------------------------------------------------------------------- 2D
def solve2d(slow, nz, nx, dz, dx, src, nsweep, eps):
    """solver2D.F90:782-978"""
    slow = np.ascontiguousarray(slow, dtype=np.float64)
    tt = np.empty(nz * nx)
    lib().or_solve2d(nz, nx, dz, dx, _d(slow), byref(src), nsweep, eps,
                     _d(tt))
    return tt


# -------------------------------------------------------------- analytic
def analytic_constant(nz, nx, ny, dz, dx, dy, z0, x0, y0, zs, xs, ys, v):
    out = np.empty(nz * nx * ny)
    lib().or_analytic_constant(nz, nx, ny, dz, dx, dy, z0, x0, y0, zs, xs, ys,
                               v, _d(out))
    return out


def analytic_gradient(nz, nx, ny, dz, dx, dy, z0, x0, y0, zs, xs, ys, vtop,
                      vbot):
    """analytic.F90:456-528; returns the x-fastest internal layout."""
    out = np.empty(nz * nx * ny)
    lib().or_analytic_gradient(nz, nx, ny, dz, dx, dy, z0, x0, y0, zs, xs, ys,
                               vtop, vbot, _d(out))
    return out


def analytic_gradient_point(nz, dz, z0, zs, xs, ys, vtop, vbot, z, x, y):
    """analytic.F90:456-528 at one point (z, x, y)."""
    f = lib().or_analytic_gradient_point
    f.argtypes = [c_int] + [c_double] * 10
    f.restype = c_double
    return f(nz, dz, z0, zs, xs, ys, vtop, vbot, z, x, y)


def xfast_to_natural(a, nz, nx, ny):
    """analytic.F90:722-731: [nz][ny][nx] x-fastest -> z-fastest."""
    return np.ascontiguousarray(
        a.reshape(nz, ny, nx).transpose(1, 2, 0)).ravel()


# --------------------------------------------------------------- locator
def padlength32(n, alignment=64):
    return lib().or_padlength32(alignment, n)


class Locator:
    """locate.F90:142-744 with the reference's state (t0Work carried over
    between events when accumulate_t0 is set)."""

    def __init__(self, nevents, ntf, ngrd):
        self.nevents, self.ntf, self.ngrd = nevents, ntf, ngrd
        self.lntf = ntf + padlength32(ntf)
        self.ldgrd = ngrd + padlength32(ngrd)
        self.tables = np.zeros(ntf * self.ldgrd, dtype=np.float32)
        self.obs = np.zeros(nevents * self.lntf, dtype=np.float32)
        self.wt = np.zeros(nevents * self.lntf, dtype=np.float32)
        self.skip = np.ones(nevents * self.lntf, dtype=np.uint8)
        self.tepoch = np.zeros(nevents)
        n = max(ngrd, 512)
        self.t0work = np.zeros(n, dtype=np.float32)
        self.objwork = np.zeros(n, dtype=np.float32)

    def set_table(self, itf, tt):
        """itf is 1-based; tt float64 -> float32 round to nearest."""
        k = (itf - 1) * self.ldgrd
        self.tables[k:k + self.ngrd] = np.asarray(tt).astype(np.float32)

    def set_observation(self, ev, obs2tf1, picks, wts):
        k = (ev - 1) * self.lntf
        obs2tf1 = np.ascontiguousarray(obs2tf1, dtype=np.int32)
        picks = np.ascontiguousarray(picks, dtype=np.float64)
        wts = np.ascontiguousarray(wts, dtype=np.float64)
        te = np.zeros(1)
        o = np.zeros(self.lntf, dtype=np.float32)
        w = np.zeros(self.lntf, dtype=np.float32)
        s = np.zeros(self.lntf, dtype=np.uint8)
        lib().or_locate_observation(picks.size, _i(obs2tf1), _d(picks),
                                    _d(wts), self.lntf, _d(te), _f(o), _f(w),
                                    s.ctypes.data_as(_bp))
        self.tepoch[ev - 1] = te[0]
        self.obs[k:k + self.lntf] = o
        self.wt[k:k + self.lntf] = w
        self.skip[k:k + self.lntf] = s

    def locate(self, ev, accumulate_t0=True, pad_argmin=True, nthreads=0):
        k = (ev - 1) * self.lntf
        o = np.ascontiguousarray(self.obs[k:k + self.lntf])
        w = np.ascontiguousarray(self.wt[k:k + self.lntf])
        s = np.ascontiguousarray(self.skip[k:k + self.lntf])
        t0 = np.zeros(1)
        ob = np.zeros(1)
        opt = lib().or_locate_event(self.ntf, self.ngrd, self.ldgrd,
                                    _f(self.tables), _f(o), _f(w),
                                    s.ctypes.data_as(_bp), self.lntf,
                                    self.tepoch[ev - 1], int(accumulate_t0),
                                    int(pad_argmin), _f(self.t0work),
                                    _f(self.objwork), nthreads, _d(t0),
                                    _d(ob))
        return opt, t0[0], ob[0]


# ------------------------------------------------- gradients and rays
def fd_gradient3d(tt, nz, nx, ny, dz, dx, dy):
    """solver3D.F90:1249-1321: [x y z] per node, natural node order."""
    t = np.ascontiguousarray(tt, dtype=np.float64)
    g = np.zeros(3 * nz * nx * ny)
    lib().or_fd_gradient3d(nz, nx, ny, c_double(dz), c_double(dx),
                           c_double(dy), _d(t), _d(g))
    return g


def fd_gradient2d(tt, nz, nx, dz, dx):
    """solver2D.F90:984-1032: (d/dx, d/dz) per node."""
    t = np.ascontiguousarray(tt, dtype=np.float64)
    gx = np.zeros(nz * nx)
    gz = np.zeros(nz * nx)
    lib().or_fd_gradient2d(nz, nx, c_double(dz), c_double(dx), _d(t), _d(gx),
                           _d(gz))
    return gx, gz


def trace_ray3d(slow, tgrad, nz, nx, ny, dz, dx, dy, integrator, src, rec,
                maxpts):
    """One ray from rec (x, y, z) to the source node src (x, y, z), both
    grid-relative metres.  Returns (ok, points [n x 3], time)."""
    pts = np.zeros(3 * maxpts)
    n = ctypes.c_int()
    t = np.zeros(1)
    rc = lib().or_trace_ray3d(nz, nx, ny, c_double(dz), c_double(dx),
                              c_double(dy), _d(slow), _d(tgrad), integrator,
                              c_double(src[0]), c_double(src[1]),
                              c_double(src[2]), c_double(rec[0]),
                              c_double(rec[1]), c_double(rec[2]), maxpts,
                              _d(pts), ctypes.byref(n), _d(t))
    return rc == 0, pts[:3 * n.value].reshape(n.value, 3), t[0]


def trace_ray2d(slow, gx, gz, nz, nx, dz, dx, integrator, src, rec, maxpts):
    """2D analogue of trace_ray3d: src, rec are (x, z); points (x, 0, z)."""
    pts = np.zeros(3 * maxpts)
    n = ctypes.c_int()
    t = np.zeros(1)
    rc = lib().or_trace_ray2d(nz, nx, c_double(dz), c_double(dx), _d(slow),
                              _d(gx), _d(gz), integrator, c_double(src[0]),
                              c_double(src[1]), c_double(rec[0]),
                              c_double(rec[1]), maxpts, _d(pts),
                              ctypes.byref(n), _d(t))
    return rc == 0, pts[:3 * n.value].reshape(n.value, 3), t[0]


# ------------------------------------------------------ reference build
_ref = None


def have_ref():
    return os.path.isfile(REF_SO)


def ref():
    """The reference Fortran built from /root/reference (oracle/Makefile).
    Loaded with lazy binding: graph3d's sort routine is not part of the
    build and is never called."""
    global _ref
    if _ref is None:
        _ref = ctypes.CDLL(REF_SO, mode=os.RTLD_LAZY)
    return _ref
