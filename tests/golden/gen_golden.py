"""Writes the golden fixtures in tests/golden/ from the reference build.

Run in the build container (needs oracle/_ref/libfteikref.so, built from the
reference sources by `make -C oracle ref`):

    python tests/golden/gen_golden.py

Every fixture holds inputs and the reference's outputs only.  The script
drives the reference's own BIND(C) entry points plus the thin forwarding
wrappers of oracle/ref_harness.f90; inputs come from a fixed seed.  Where an
input state is "realistic" (neighbour times taken from a partly solved
field), that state is produced by the oracle, but the expected outputs are
always the reference's.
"""
import ctypes
import os
import sys
from ctypes import POINTER, byref, c_bool, c_double, c_float, c_int

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle as orc  # noqa: E402

_dp = POINTER(c_double)
_fp = POINTER(c_float)
_ip = POINTER(c_int)


def D(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(_dp)


def F(a):
    return a.ctypes.data_as(_fp)


def I(a):
    return a.ctypes.data_as(_ip)


R = orc.ref()
ierr = c_int(0)


def chk(msg):
    if ierr.value != 0:
        raise RuntimeError("reference call failed: " + msg)


# --------------------------------------------------------------- helpers
from tests.golden.gen_golden_models import hetero_model3d, smooth_noise  # noqa: E402,E501


def set_model3d(nz, nx, ny, dz, dx, dy, z0, x0, y0, vel_cells):
    R.fteik_model_initializeGeometry(c_bool(True), c_int(nz), c_int(nx),
                                     c_int(ny), c_double(z0), c_double(x0),
                                     c_double(y0), c_double(dz), c_double(dx),
                                     c_double(dy), byref(ierr))
    chk("initializeGeometry")
    R.fteik_model_setVelocityModel64f(c_int(vel_cells.size), D(vel_cells),
                                      byref(ierr))
    chk("setVelocityModel64f")


def set_source(zs, xs, ys):
    z = np.array([zs]); x = np.array([xs]); y = np.array([ys])
    R.fteik_source_initialize64f(c_int(1), D(z), D(x), D(y), c_int(0),
                                 byref(ierr))
    chk("source_initialize")


def source_info(isrc=1):
    zsi, xsi, ysi = c_int(), c_int(), c_int()
    zsa, xsa, ysa, sz, sz2 = (c_double() for _ in range(5))
    R.fteik_source_getSolverInfo64fF(c_int(isrc), byref(zsi), byref(xsi),
                                     byref(ysi), byref(zsa), byref(xsa),
                                     byref(ysa), byref(sz), byref(sz2),
                                     byref(ierr))
    chk("getSolverInfo")
    return (zsi.value, xsi.value, ysi.value, zsa.value, xsa.value, ysa.value,
            sz.value)


out = {}

# ------------------------------------------------- 3D model / source
rng = np.random.default_rng(20240601)
nz, nx, ny, h = 17, 13, 11, 25.0
vel = hetero_model3d(rng, nz, nx, ny, h)
set_model3d(nz, nx, ny, h, h, h, 0.0, 0.0, 0.0, vel)
# cell model orders 1 and 2 and nodal model orders 0,1,2: velocity readback
vel_back = {}
for kind, order in (("cell", 1), ("cell", 2), ("nodal", 0), ("nodal", 1),
                    ("nodal", 2)):
    if kind == "cell":
        v = 2000.0 + 1000.0 * rng.random((nz - 1) * (nx - 1) * (ny - 1))
        R.fteik_model_setCellVelocityModel64f(c_int(v.size), c_int(order),
                                              D(v), byref(ierr))
    else:
        v = 2000.0 + 1000.0 * rng.random(nz * nx * ny)
        R.fteik_model_setNodalVelocityModel64f(c_int(v.size), c_int(order),
                                               D(v), byref(ierr))
    chk("set %s %d" % (kind, order))
    nv = c_int()
    back = np.zeros((nz - 1) * (nx - 1) * (ny - 1))
    R.fteik_model_getVelocityModel64f(c_int(back.size), byref(nv), D(back),
                                      byref(ierr))
    chk("getVelocityModel")
    out["model_%s%d_in" % (kind, order)] = v
    out["model_%s%d_velback" % (kind, order)] = back
out["model_grid"] = np.array([nz, nx, ny], dtype=np.int32)

# sources: interior, on faces, on the far face (nudged)
set_model3d(nz, nx, ny, h, h, h, 10.0, -5.0, 3.0, vel)
srcs = np.array([[10.0 + 7.3 * h, -5.0 + 4.0 * h, 3.0 + 5.5 * h],
                 [10.0, -5.0 + 3.2 * h, 3.0],
                 [10.0 + (nz - 1) * h, -5.0 + (nx - 1) * h, 3.0 + 2 * h],
                 [10.0 + 3 * h, -5.0 + 6 * h, 3.0 + 4 * h]])
info = []
for s in srcs:
    set_source(*s)
    info.append(source_info())
out["source_grid"] = np.array([nz, nx, ny, h, h, h, 10.0, -5.0, 3.0])
out["source_in"] = srcs
out["source_ints"] = np.array([i[:3] for i in info], dtype=np.int32)
out["source_reals"] = np.array([i[3:] for i in info])

# ------------------------------------------------------- receivers
tt = rng.random(nz * nx * ny) * 3.0
zr = np.concatenate([rng.uniform(0, (nz - 1) * h, 20), [0.0, (nz - 1) * h,
                                                         -3.0]])
xr = np.concatenate([rng.uniform(0, (nx - 1) * h, 20), [(nx - 1) * h, 0.0,
                                                        5.0]])
yr = np.concatenate([rng.uniform(0, (ny - 1) * h, 20), [0.0, (ny - 1) * h,
                                                        7.0]])
set_model3d(nz, nx, ny, h, h, h, 0.0, 0.0, 0.0, vel)
R.fteik_receiver_initialize64f(c_int(zr.size), D(zr), D(xr), D(yr), c_int(0),
                               byref(ierr))
rec_ierr = ierr.value
ttr = np.zeros(zr.size)
R.fteik_receiver_getTravelTimes64f(c_int(zr.size), c_int(tt.size), D(tt),
                                   D(ttr), byref(ierr))
out["rec_field"] = tt
out["rec_zxy"] = np.stack([zr, xr, yr])
out["rec_ttr"] = ttr
out["rec_ierr"] = np.array([rec_ierr])

# ---------------------------------------------------------- gathers
slow = 1.0 / vel
ttf = rng.random(nz * nx * ny) * 2.0
gz, gs = [], []
for sweep in range(1, 9):
    lim = orc.sweep_limits(sweep, 0, nz, nx, ny, 1, 1, 1)
    for _ in range(40):
        i = int(rng.integers(min(lim[0], lim[1]), max(lim[0], lim[1]) + 1))
        j = int(rng.integers(min(lim[2], lim[3]), max(lim[2], lim[3]) + 1))
        k = int(rng.integers(min(lim[4], lim[5]), max(lim[4], lim[5]) + 1))
        s7 = [c_double() for _ in range(7)]
        fn = getattr(R, "fteik_prefetchSweep%dSlowness64fF" % sweep)
        ci = [c_int(v) for v in (i, j, k, nz, nx, ny, nz - 1,
                                 (nz - 1) * (nx - 1))]
        fn(*[byref(v) for v in ci], D(slow), *[byref(v) for v in s7])
        t8 = [c_double() for _ in range(8)]
        fn = getattr(R, "fteik_prefetchSweep%dTravelTimes64fF" % sweep)
        fn(c_int(i), c_int(j), c_int(k), c_int(nz), c_int(nx), c_int(ny),
           c_int(nz * nx), D(ttf), *[byref(v) for v in t8])
        gz.append([sweep, i, j, k])
        gs.append([v.value for v in s7] + [v.value for v in t8])
out["gather_slow"] = slow
out["gather_tt"] = ttf
out["gather_nodes"] = np.array(gz, dtype=np.int32)
out["gather_vals"] = np.array(gs)

# ------------------------------------------------- 3D local solvers
# Realistic neighbour states: the oracle's field after the init pass, plus
# random perturbations, evaluated by the reference's stencils.
nz, nx, ny, h = 21, 19, 17, 20.0
vel = hetero_model3d(rng, nz, nx, ny, h)
slow = 1.0 / vel
zs, xs, ys = 8.3 * h, 9.6 * h, 7.2 * h
eps = 3.0
set_model3d(nz, nx, ny, h, h, h, 0.0, 0.0, 0.0, vel)
set_source(zs, xs, ys)
dest = np.zeros(8, dtype=np.int32)
ts8 = np.zeros(8)
R.refh_ls3d_initialize(c_int(1), c_double(eps), I(dest), D(ts8), byref(ierr))
chk("ls3d_initialize")
out["ls3d_grid"] = np.array([nz, nx, ny, h, zs, xs, ys, eps])
out["ls3d_vel"] = vel
out["ls3d_dest"] = dest
out["ls3d_ts8"] = ts8
src = orc.source_locate((nz, nx, ny, h, h, h, 0.0, 0.0, 0.0), zs, xs, ys,
                        slow)
partial = orc.solve3d(slow, nz, nx, ny, h, h, h, src, 0, eps)
zsi, xsi, ysi = src.zsi, src.xsi, src.ysi
rows_in, rows_out_noinit, rows_out_init, rows_node = [], [], [], []
for trial in range(1000):
    sweep = int(rng.integers(1, 9))
    near = trial % 2 == 0
    lim = orc.sweep_limits(sweep, 0, nz, nx, ny, zsi, xsi, ysi)
    lo = [min(lim[0], lim[1]), min(lim[2], lim[3]), min(lim[4], lim[5])]
    hi = [max(lim[0], lim[1]), max(lim[2], lim[3]), max(lim[4], lim[5])]
    if near:
        c = (zsi, xsi, ysi)
        ijk = [int(np.clip(c[a] + rng.integers(-3, 5), lo[a], hi[a]))
               for a in range(3)]
    else:
        ijk = [int(rng.integers(lo[a], hi[a] + 1)) for a in range(3)]
    i, j, k = ijk
    s7 = orc.gather_slow3d(sweep, i, j, k, nz, nx, ny, slow)
    t7, t8 = orc.gather_tt3d(sweep, i, j, k, nz, nx, partial)
    mode = trial % 5
    if mode == 1:
        t7 = t7 * (1.0 + 0.02 * rng.standard_normal(7))
    elif mode == 2:
        m = rng.random(7) < 0.3
        t7 = np.where(m, orc.HUGE, t7)
    elif mode == 3:
        t8 = t8 * (1.0 + 0.05 * rng.standard_normal())
    elif mode == 4:
        s7 = s7 * (1.0 + 0.1 * rng.standard_normal(7))
    sgn = [(1, 1, 1), (1, -1, 1), (1, 1, -1), (1, -1, -1), (-1, 1, 1),
           (-1, -1, 1), (-1, 1, -1), (-1, -1, -1)][sweep - 1]
    tup = np.array([t8])
    a = [np.array([v]) for v in t7]
    b = [np.array([v]) for v in s7]
    R.refh_ls3d_noinit(c_int(1), *[D(v) for v in a], *[D(v) for v in b],
                       D(tup))
    R.refh_ls3d_init.restype = c_double
    ti = R.refh_ls3d_init(*[c_double(v) for v in t7], c_double(t8),
                          *[c_double(v) for v in s7], c_int(i), c_int(j),
                          c_int(k), c_int(sgn[0]), c_int(sgn[1]),
                          c_int(sgn[2]), c_double(sgn[0] / h),
                          c_double(sgn[1] / h), c_double(sgn[2] / h))
    rows_in.append(list(t7) + [t8] + list(s7))
    rows_node.append([sweep, i, j, k])
    rows_out_noinit.append(tup[0])
    rows_out_init.append(ti)
out["ls3d_in"] = np.array(rows_in)
out["ls3d_node"] = np.array(rows_node, dtype=np.int32)
out["ls3d_noinit"] = np.array(rows_out_noinit)
out["ls3d_init"] = np.array(rows_out_init)


# --------------------------------------------------------------- 2D
def solve2d_fsm_ref(nz, nx, dz, dx, vel, zs, xs, nsweep, eps):
    """solver2D.F90:1042-1280 (the fast-sweeping-method driver)."""
    R.fteik_solver2d_initialize(c_int(nz), c_int(nx), c_double(0.0),
                                c_double(0.0), c_double(dz), c_double(dx),
                                c_int(nsweep), c_double(eps), c_double(0.0),
                                c_int(0), byref(ierr))
    chk("solver2d_initialize (FSM)")
    R.fteik_solver2d_setCellVelocityModel64f(c_int(vel.size), c_int(0),
                                             D(vel), byref(ierr))
    chk("solver2d vel (FSM)")
    R.fteik_solver2d_setSources64f(c_int(1), D(np.array([zs])),
                                   D(np.array([xs])), byref(ierr))
    chk("solver2d src (FSM)")
    R.fteik_solver2d_solveSourceFSM(c_int(1), byref(ierr))
    chk("solver2d FSM solve")
    tt = np.zeros(nz * nx)
    R.fteik_solver2d_getTravelTimeField64f(c_int(tt.size), c_int(0), D(tt),
                                           byref(ierr))
    chk("solver2d FSM field")
    R.fteik_solver2d_free()
    return tt


def gradient2d_ref(nz, nx, dz, dx, vel, zs, xs, nsweep, eps):
    """solver2D.F90:984-1032 on the reference's solved field (through
    refh_solver2d_gradient, which reaches it by the reference's own
    traceRaysToPoints path)."""
    R.fteik_solver2d_initialize(c_int(nz), c_int(nx), c_double(0.0),
                                c_double(0.0), c_double(dz), c_double(dx),
                                c_int(nsweep), c_double(eps), c_double(0.0),
                                c_int(0), byref(ierr))
    chk("solver2d_initialize (gradient)")
    R.fteik_solver2d_setCellVelocityModel64f(c_int(vel.size), c_int(0),
                                             D(vel), byref(ierr))
    R.fteik_solver2d_setSources64f(c_int(1), D(np.array([zs])),
                                   D(np.array([xs])), byref(ierr))
    R.fteik_solver2d_solveSourceLSM(c_int(1), byref(ierr))
    chk("solver2d solve (gradient)")
    gx = np.zeros(nz * nx)
    gz = np.zeros(nz * nx)
    e = c_int(0)
    R.refh_solver2d_gradient(c_int(nz * nx), D(gx), D(gz), byref(e))
    # the out-of-model destination makes the call itself fail after the
    # gradient is computed
    assert e.value == 1
    R.fteik_solver2d_free()
    return gx, gz


def solve2d_ref(nz, nx, dz, dx, vel, zs, xs, nsweep, eps, zr, xr):
    R.fteik_solver2d_initialize(c_int(nz), c_int(nx), c_double(0.0),
                                c_double(0.0), c_double(dz), c_double(dx),
                                c_int(nsweep), c_double(eps), c_double(0.0),
                                c_int(0), byref(ierr))
    chk("solver2d_initialize")
    R.fteik_solver2d_setCellVelocityModel64f(c_int(vel.size), c_int(0),
                                             D(vel), byref(ierr))
    chk("solver2d vel")
    R.fteik_solver2d_setSources64f(c_int(1), D(np.array([zs])),
                                   D(np.array([xs])), byref(ierr))
    chk("solver2d src")
    R.fteik_solver2d_setReceivers64f(c_int(zr.size), D(zr), D(xr),
                                     byref(ierr))
    chk("solver2d rec")
    R.fteik_solver2d_solveSourceLSM(c_int(1), byref(ierr))
    chk("solver2d solve")
    tt = np.zeros(nz * nx)
    R.fteik_solver2d_getTravelTimeField64f(c_int(tt.size), c_int(0), D(tt),
                                           byref(ierr))
    chk("solver2d field")
    ttr = np.zeros(zr.size)
    R.fteik_solver2d_getTravelTimes64f(c_int(zr.size), D(ttr), byref(ierr))
    chk("solver2d ttr")
    ttc = np.zeros((nz - 1) * (nx - 1))
    R.fteik_solver2d_getTravelTimeFieldCell64f(c_int(ttc.size), c_int(1),
                                               D(ttc), byref(ierr))
    chk("solver2d cell")
    R.fteik_solver2d_free()
    return tt, ttr, ttc


# reference test: testing/test2d.F90:12-93 (201 x 101, v 3000 -> 5000)
nx2, nz2 = 201, 101
x1, z1 = 5.0e3, 2.5e3
dx2, dz2 = x1 / (nx2 - 1), z1 / (nz2 - 1)
vel2 = np.zeros((nz2 - 1) * (nx2 - 1))
for ix in range(1, nx2):
    for iz in range(1, nz2):
        z = (iz - 1) * dz2
        vel2[(ix - 1) * (nz2 - 1) + iz - 1] = 3000.0 + (5000.0 - 3000.0) / z1 * z
zs2, xs2 = 2.0 * z1 / 3.0, x1 / 3.0
zr2 = np.array([0.0, 100.0, 1250.0, 2499.0])
xr2 = np.array([0.0, 4999.0, 2500.0, 10.0])
tt2, ttr2, ttc2 = solve2d_ref(nz2, nx2, dz2, dx2, vel2, zs2, xs2, 2, 5.0, zr2,
                              xr2)
orig = np.zeros(nz2 * nx2)
slow2 = np.ascontiguousarray((1.0 / vel2))
R.refh_orig_solver2d(D(slow2), D(orig), c_int(nz2), c_int(nx2),
                     c_double(zs2), c_double(xs2), c_double(dz2),
                     c_double(dx2), c_int(2))
out["t2d_a_cfg"] = np.array([nz2, nx2, dz2, dx2, zs2, xs2, 2, 5.0])
out["t2d_a_tt"] = tt2
out["t2d_a_rec"] = np.stack([zr2, xr2])
out["t2d_a_ttr"] = ttr2
# testing/test2d.F90:84-87 asserts max|LSM - original| <= 1e-10
out["t2d_a_orig_maxdiff"] = np.array([np.abs(orig - tt2).max()])
out["t2d_a_orig_sample"] = orig[::97]

# heterogeneous small 2D case with a source near the origin corner
nz3, nx3, h3 = 41, 37, 15.0
g = smooth_noise(rng, (nx3 - 1, nz3 - 1), 2.0)
vel3 = ((2000.0 + 3.0 * (np.arange(nz3 - 1) + 0.5) * h3)[None, :]
        * (1.0 + 0.1 * g)).ravel()
zs3, xs3 = 2.4 * h3, 30.7 * h3
zr3 = rng.uniform(0, (nz3 - 1) * h3, 12)
xr3 = rng.uniform(0, (nx3 - 1) * h3, 12)
tt3, ttr3, ttc3 = solve2d_ref(nz3, nx3, h3, h3, vel3, zs3, xs3, 3, 4.0, zr3,
                              xr3)
out["t2d_b_cfg"] = np.array([nz3, nx3, h3, h3, zs3, xs3, 3, 4.0])
out["t2d_b_vel"] = vel3
out["t2d_b_tt"] = tt3
out["t2d_b_rec"] = np.stack([zr3, xr3])
out["t2d_b_ttr"] = ttr3
out["t2d_b_ttcell_xz"] = ttc3
# the finite-difference gradient of the same two solved fields
out["t2d_a_gradx"], out["t2d_a_gradz"] = gradient2d_ref(
    nz2, nx2, dz2, dx2, vel2, zs2, xs2, 2, 5.0)
out["t2d_b_gradx"], out["t2d_b_gradz"] = gradient2d_ref(
    nz3, nx3, h3, h3, vel3, zs3, xs3, 3, 4.0)
# the FSM driver on the same two configurations
out["t2d_a_fsm_tt"] = solve2d_fsm_ref(nz2, nx2, dz2, dx2, vel2, zs2, xs2, 2,
                                      5.0)
out["t2d_b_fsm_tt"] = solve2d_fsm_ref(nz3, nx3, h3, h3, vel3, zs3, xs3, 3,
                                      4.0)

# 2D stencils on random states
nz4, nx4 = 15, 12
R.fteik_solver2d_initialize(c_int(nz4), c_int(nx4), c_double(0.0),
                            c_double(0.0), c_double(12.0), c_double(9.0),
                            c_int(1), c_double(3.0), c_double(0.0), c_int(0),
                            byref(ierr))
vel4 = 1500.0 + 2000.0 * rng.random((nz4 - 1) * (nx4 - 1))
R.fteik_solver2d_setCellVelocityModel64f(c_int(vel4.size), c_int(0), D(vel4),
                                         byref(ierr))
R.fteik_solver2d_setSources64f(c_int(1), D(np.array([4.4 * 12.0])),
                               D(np.array([5.1 * 9.0])), byref(ierr))
d4 = np.zeros(4, dtype=np.int32)
ts4 = np.zeros(4)
R.refh_ls2d_initialize(c_int(1), c_double(3.0), I(d4), D(ts4), byref(ierr))
chk("ls2d_initialize")
slow4 = 1.0 / vel4
base = np.abs(np.indices((nx4, nz4))[0] - 5.1) * 9.0 * 4e-4 + \
    np.abs(np.indices((nx4, nz4))[1] - 4.4) * 12.0 * 4e-4
rows = []
for trial in range(400):
    sweep = int(rng.integers(1, 5))
    sgntz = 1 if sweep <= 2 else -1
    sgntx = 1 if sweep in (1, 3) else -1
    iz = int(rng.integers(2, nz4)) if sgntz > 0 else int(rng.integers(1, nz4))
    ix = int(rng.integers(2, nx4)) if sgntx > 0 else int(rng.integers(1, nx4))
    ttf2 = (base.ravel() * (1.0 + 0.05 * rng.standard_normal(nz4 * nx4)))
    if trial % 3 == 0:
        ttf2[rng.random(nz4 * nx4) < 0.2] = orc.HUGE
    tup = c_double()
    R.refh_ls2d_init(c_int(sweep), c_int(nz4), c_int(nx4), c_int(iz),
                     c_int(ix), D(slow4), D(ttf2), byref(tup))
    tv = ttf2[(ix - 1) * nz4 + iz - sgntz - 1]
    te = ttf2[(ix - sgntx - 1) * nz4 + iz - 1]
    tev = ttf2[(ix - sgntx - 1) * nz4 + iz - sgntz - 1]
    t0 = ttf2[(ix - 1) * nz4 + iz - 1]
    sl = rng.random(3) * 5e-4 + 2e-4
    ttvec = np.array([tv, te, tev, t0])
    sloc = np.array([sl[0], sl[1], sl[2], 0.0])
    tn = np.zeros(1)
    R.refh_ls2d_noinit(c_int(1), D(ttvec), D(sloc), D(tn))
    # init64f reads only tv, te, tev, tt from the field (localSolver.F90:
    # 213-220), so the row keeps those four values.
    rows.append([sweep, iz, ix, tup.value, tn[0]] + list(ttvec) + list(sl))
R.fteik_solver2d_free()
out["ls2d_cfg"] = np.array([nz4, nx4, 12.0, 9.0, 4.4 * 12.0, 5.1 * 9.0, 3.0])
out["ls2d_vel"] = vel4
out["ls2d_dest"] = d4
out["ls2d_ts4"] = ts4
out["ls2d_rows"] = np.array(rows)

# ----------------------------------------------------------- analytic
na = (11, 13, 9)
R.fteik_analytic_initialize64f(c_int(na[0]), c_int(na[1]), c_int(na[2]),
                               c_double(0.0), c_double(0.0), c_double(0.0),
                               c_double(10.0), c_double(12.0), c_double(8.0),
                               c_int(0), byref(ierr))
chk("analytic init")
R.fteik_analytic_setSources64f(c_int(1), D(np.array([42.0])),
                               D(np.array([61.0])), D(np.array([33.0])),
                               byref(ierr))
chk("analytic src")
R.fteik_analytic_setLinearVelocityGradient64f(c_double(3000.0),
                                              c_double(3500.0), byref(ierr))
chk("analytic grad")
R.fteik_analytic_solveSourceLinearVelocityGradient64f(c_int(1), byref(ierr))
chk("analytic solve")
nga = na[0] * na[1] * na[2]
fields = {}
for order in (0, 1, 2):
    f = np.zeros(nga)
    R.fteik_analytic_getTravelTimeField64f(c_int(nga), c_int(order), D(f),
                                           byref(ierr))
    fields[order] = f
out["ana_cfg"] = np.array([na[0], na[1], na[2], 10.0, 12.0, 8.0, 42.0, 61.0,
                           33.0, 3000.0, 3500.0])
out["ana_grad_natural"] = fields[0]
out["ana_grad_xyz"] = fields[1]
out["ana_grad_zyx"] = fields[2]
R.fteik_analytic_setConstantVelocity64f(c_double(4200.0), byref(ierr))
R.fteik_analytic_solveSourceConstantVelocity64f(c_int(1), byref(ierr))
chk("analytic const")
f = np.zeros(nga)
R.fteik_analytic_getTravelTimeField64f(c_int(nga), c_int(0), D(f),
                                       byref(ierr))
out["ana_const_natural"] = f
R.fteik_analytic_free()

# ------------------------------------------------------------ locator
def locator_case(tag, nz, nx, ny, h, nsta, nev, seed):
    r = np.random.default_rng(seed)
    ngrd = nz * nx * ny
    R.locate_initialize(c_int(nev), c_int(nsta), c_int(ngrd), byref(ierr))
    chk("locate_initialize")
    zz, xx, yy = np.meshgrid(np.arange(nz) * h, np.arange(nx) * h,
                             np.arange(ny) * h, indexing="ij")
    # natural ordering (z fastest): index (iy*nx + ix)*nz + iz
    zz = zz.transpose(2, 1, 0).ravel()
    xx = xx.transpose(2, 1, 0).ravel()
    yy = yy.transpose(2, 1, 0).ravel()
    sx = r.uniform(0, (nx - 1) * h, nsta)
    sy = r.uniform(0, (ny - 1) * h, nsta)
    tabs = []
    for s in range(nsta):
        v = 5000.0 if s < nsta - 1 else 5000.0 / np.sqrt(3.0)
        t = np.sqrt(zz ** 2 + (xx - sx[s]) ** 2 + (yy - sy[s]) ** 2) / v
        tabs.append(t)
        R.locate_setTravelTimeField64f(c_int(ngrd), c_int(s + 1), D(t),
                                       byref(ierr))
        chk("setTravelTimeField64f")
    true_idx = r.integers(0, ngrd, nev)
    ot = r.uniform(0, 100, nev)
    picks_all, wts_all, res = [], [], []
    for e in range(nev):
        picks = np.array([tabs[s][true_idx[e]] for s in range(nsta)]) + ot[e]
        picks = picks + r.normal(0.0, 0.02, nsta) * (e % 2)
        wts = np.ones(nsta)
        if e % 3 == 2:
            wts[r.integers(0, nsta)] = 0.0
            wts[r.integers(0, nsta)] = 0.5
        o2t = np.arange(1, nsta + 1, dtype=np.int32)
        R.locate_setObservation64f(c_int(e + 1), c_int(nsta), c_bool(False),
                                   c_double(0.0), I(o2t), D(picks), D(wts),
                                   byref(ierr))
        chk("setObservation")
        picks_all.append(picks)
        wts_all.append(wts)
    for e in range(nev):
        opt, t0, ob = c_int(), c_double(), c_double()
        R.locate_locateEvent(c_int(e + 1), byref(opt), byref(t0), byref(ob))
        res.append([opt.value, t0.value, ob.value])
    R.locate_finalize()
    out["loc_%s_cfg" % tag] = np.array([nz, nx, ny, h, nsta, nev])
    out["loc_%s_tables" % tag] = np.array(tabs)
    out["loc_%s_picks" % tag] = np.array(picks_all)
    out["loc_%s_wts" % tag] = np.array(wts_all)
    out["loc_%s_true" % tag] = true_idx
    out["loc_%s_result" % tag] = np.array(res)


# 9 x 8 x 7 = 504 < 512: exercises the padded MINLOC (locate.F90:200-203)
locator_case("small", 9, 8, 7, 100.0, 5, 10, 7)
locator_case("mid", 15, 13, 11, 100.0, 6, 12, 11)

np.savez_compressed(os.path.join(HERE, "fixtures.npz"), **out)
print("wrote", os.path.join(HERE, "fixtures.npz"),
      os.path.getsize(os.path.join(HERE, "fixtures.npz")), "bytes",
      len(out), "arrays")
