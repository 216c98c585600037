"""Benchmark: 3D level-set fast sweeping at 257^3 on one MI355X.

Workload: grid nz = nx = ny = 257, h = 20 m, 8 flat layers at seeded random
depths with velocities 1500 -> 6500 m/s increasing with depth, times a
Gaussian-smoothed (sigma = 5 cells) +-5 % random perturbation; one source
at a seeded random position; the initialization sweeps, then as many
8-sweep Gauss-Seidel iterations as the FP64 field needs to stop changing
(found at start-up and reported as nsweep).

One step = one complete single-source solve (fteik_solver3d_solveSourceLSM)
with the model and source already resident on the GPU.  The FP64 path
(reference semantics, bitwise equal to the oracle) and the mixed path (FP32
Gauss-Seidel iterations, then FP64 correction iterations to convergence) are
each timed over K steps after W warm-up steps.

Output: one JSON line.  value = node updates per second of the FP64 solve
(updates as the reference counts them: initialization masks plus
nsweep x 8 x (n-1)^3), in Gupdates/s, for the dense sweeps (every update
performed).  The same solve with tile skipping and with chained sweeps is
reported as time to the same field.  The mixed path runs nsweep FP32
iterations and FP64 iterations until one changes nothing; its rate on the
same update count, its speed-up and its error against the FP64 field are
reported beside it.  roofline: the FP64 Gauss-Seidel sweep kernel, timed
with hipEvents on the library's stream; cpu_baseline: the oracle's OpenMP
level-synchronous FP64 solve of the same workload on the host cores.
"""
import argparse
import ctypes
import json
import os
import sys
import time
from ctypes import byref, c_double, c_int

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N, H = 257, 20.0
BYTES_PER_UPDATE = {"f64": 24.0, "f32": 12.0}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "sweep3d_257_traffic.json")


def layered_model(n=N, h=H):
    """Cell velocities (natural order, z fastest) and a source position:
    7 interface depths sorted U(0, (n-1) h) (seed 2), layer velocities
    1500 -> 6500 m/s from top to bottom, times 1 + 0.05 g with g a
    Gaussian-smoothed (sigma = 5 cells) N(0,1) field normalised to
    max|g| = 1 (seed 3); source uniform over the cells whose indices are
    all <= n-3 (0-based, seed 4)."""
    from scipy.ndimage import gaussian_filter
    nc = n - 1
    interfaces = np.sort(np.random.default_rng(2).uniform(0.0, nc * h, 7))
    vlayer = np.linspace(1500.0, 6500.0, 8)
    zc = (np.arange(nc) + 0.5) * h
    v1d = vlayer[np.searchsorted(interfaces, zc)]
    g = gaussian_filter(np.random.default_rng(3).standard_normal((nc, nc, nc)),
                        sigma=5.0, mode="reflect")
    g /= np.abs(g).max()
    vel = v1d[None, None, :] * (1.0 + 0.05 * g)
    src = np.random.default_rng(4).uniform(0.0, (n - 2) * h, 3)  # (z, x, y)
    return np.ascontiguousarray(vel.ravel()), src


def batch_models(ngrids=64, h=50.0):
    """64 grids, each axis U{64..192} (seed 5); v = 2000 + U(0.3,0.8) z
    (seed 6) times 1 + 0.03 g, g smoothed as for the 257^3 model (seed 7);
    one source per grid at a random node with every index <= n-3 (seed 8)."""
    from scipy.ndimage import gaussian_filter
    dims = np.random.default_rng(5).integers(64, 193, (ngrids, 3))
    grad = np.random.default_rng(6).uniform(0.3, 0.8, ngrids)
    r7 = np.random.default_rng(7)
    r8 = np.random.default_rng(8)
    out = []
    for k in range(ngrids):
        nz, nx, ny = (int(v) for v in dims[k])
        zc = (np.arange(nz - 1) + 0.5) * h
        g = gaussian_filter(r7.standard_normal((ny - 1, nx - 1, nz - 1)),
                            sigma=5.0, mode="reflect")
        g /= np.abs(g).max()
        vel = (2000.0 + grad[k] * zc)[None, None, :] * (1.0 + 0.03 * g)
        node = [int(r8.integers(0, n - 2)) for n in (nz, nx, ny)]
        out.append(((nz, nx, ny), np.ascontiguousarray(vel.ravel()),
                    tuple(i * h for i in node)))
    return out


def batch_setup(L, models, h, precision):
    ierr = c_int(1)
    dp = ctypes.POINTER(c_double)
    L.fteik_batch3d_initialize(len(models), byref(ierr))
    assert ierr.value == 0
    for k, ((nz, nx, ny), vel, (zs, xs, ys)) in enumerate(models):
        L.fteik_batch3d_setGrid(k + 1, nz, nx, ny, 0.0, 0.0, 0.0, h, h, h,
                                vel.ctypes.data_as(dp), zs, xs, ys,
                                byref(ierr))
        assert ierr.value == 0, "batch grid %d" % k
    L.fteik_batch3d_setPrecision(precision, byref(ierr))
    L.fteik_batch3d_setMixedParameters(1.0 / 64.0, 64, byref(ierr))
    assert ierr.value == 0


def batch_info(L):
    ms, upd, corr, ierr = c_double(), c_double(), c_int(), c_int()
    L.fteik_batch3d_getLastSolveInfo(byref(ms), byref(corr), byref(upd),
                                     byref(ierr))
    return ms.value, upd.value, corr.value


def batch_fields(L, models):
    ierr = c_int(1)
    out = []
    for k, ((nz, nx, ny), _, _) in enumerate(models):
        tt = np.zeros(nz * nx * ny)
        L.fteik_batch3d_getTravelTimeField64f(
            k + 1, tt.size, 0, tt.ctypes.data_as(ctypes.POINTER(c_double)),
            byref(ierr))
        assert ierr.value == 0
        out.append(tt)
    return out


def batch_section(L, args):
    """Batched solve of 64 independent grids of uneven size in FP32-mixed
    precision (FP32 iterations, then FP64 corrections to convergence), with
    the tile shape chosen by the library's autotuner on the batch.  The
    FP64 batch is solved to convergence too (the iteration count after
    which one more changes no node): it is the parity reference of the
    mixed fields (max relative error per grid), and a seeded sample of its
    grids at nsweep = 2 is checked bitwise against the oracle."""
    h = 50.0
    models = batch_models()
    ierr = c_int(1)
    # FP64 iterations to convergence (default tile)
    prev, nconv = None, None
    for ns in range(2, 33, 2):
        batch_setup(L, models, h, 0)
        L.fteik_batch3d_solve(ns, 5.0, byref(ierr))
        assert ierr.value == 0
        cur = batch_fields(L, models)
        if prev is not None and all(np.array_equal(a, b)
                                    for a, b in zip(cur, prev)):
            nconv = ns - 2
            break
        prev = cur
        L.fteik_batch3d_free()
    if nconv is None:
        nconv = 32
    conv = prev
    # FP64 at nconv, autotuned, timed
    batch_setup(L, models, h, 0)
    t64 = c_int(-1)
    L.fteik_batch3d_autotune(nconv, 5.0, byref(t64), byref(ierr))
    assert ierr.value == 0, "batch autotune (FP64)"
    ms64, upd64, _ = batch_info(L)
    L.fteik_batch3d_free()
    # FP32-mixed, autotuned, timed
    batch_setup(L, models, h, 1)
    tmx = c_int(-1)
    L.fteik_batch3d_autotune(nconv, 5.0, byref(tmx), byref(ierr))
    assert ierr.value == 0, "batch autotune (mixed)"
    L.fteik_batch3d_solve(nconv, 5.0, byref(ierr))
    assert ierr.value == 0, "batch mixed solve"
    msmx, updmx, corr = batch_info(L)
    mixed = batch_fields(L, models)
    rel = [float(np.max(np.abs(m - c) / np.maximum(c, 1e-12)))
           for m, c in zip(mixed, conv)]
    L.fteik_batch3d_free()
    out = {"workload": "64 grids, each axis U{64..192}, h=50 m, "
                       "v=2000+U(0.3,0.8)z, +-3% smoothed, one source each, "
                       "FP32-mixed, autotuned tile",
           "grids": len(models), "fp64_iterations_to_converge": nconv,
           "value": upd64 / (msmx * 1e-3) / 1e9, "unit": "Gupdates/s",
           "ms": msmx, "autotuned_tile": tmx.value,
           "fp64_correction_iterations": corr,
           "updates_performed": updmx,
           "mixed_max_rel_err_vs_fp64_per_grid": max(rel),
           "grids_within_1e-6": int(sum(r <= 1e-6 for r in rel)),
           "fp64": {"value": upd64 / (ms64 * 1e-3) / 1e9, "ms": ms64,
                    "autotuned_tile": t64.value,
                    "updates": upd64},
           "speedup_mixed_vs_fp64": ms64 / msmx}
    if not args.no_cpu:
        from oracle import oracle as orc
        batch_setup(L, models, h, 0)
        L.fteik_batch3d_solve(2, 5.0, byref(ierr))
        f2 = batch_fields(L, models)
        L.fteik_batch3d_free()
        same = True
        for k in np.random.default_rng(55).choice(len(models), 2,
                                                  replace=False):
            (nz, nx, ny), vel, (zs, xs, ys) = models[k]
            slow = orc.cell_slowness(vel, nz, nx, ny)
            src = orc.source_locate((nz, nx, ny, h, h, h, 0.0, 0.0, 0.0), zs,
                                    xs, ys, slow)
            ref = orc.solve3d(slow, nz, nx, ny, h, h, h, src, 2, 5.0,
                              schedule=1)
            same = same and bool(np.array_equal(ref, f2[k]))
        out["fp64_sample_equals_oracle"] = same
    return out


def locator_workload():
    """Grid-search workload: 161 x 161 x 81 (nx = ny = 161, nz = 81),
    h = 250 m, v = 4000 + 0.4 z; 48 stations at z = 0, x, y ~ U[0, 158 h)
    (seed 9); 10,000 events at uniform random nodes (seed 10), origin
    times U(0, 100) s (seed 11), picks = T + origin + N(0, 0.05^2)
    (seed 12), weights 1."""
    nz, nx, ny, h = 81, 161, 161, 250.0
    zc = (np.arange(nz - 1) + 0.5) * h
    vel = np.ascontiguousarray(np.tile(4000.0 + 0.4 * zc,
                                       (nx - 1) * (ny - 1)))
    st = np.random.default_rng(9).uniform(0.0, 158.0 * h, (48, 2))
    stations = [(0.0, float(x), float(y)) for x, y in st]
    return (nz, nx, ny, h), vel, stations


def solver2d_section(L, nsrc=256, cpu_sources=4):
    """2D solver, many sources in one launch (one workgroup per source):
    501 x 501 nodes, h = 10 m, v = 1500 + 0.8 z times a +-5 % smooth
    perturbation; 256 sources at seeded random nodes (seed 21), 32 surface
    receivers; nsweep = 2, eps = 3 cells.  Parity: the receiver times of a
    sample of sources equal the oracle's; CPU baseline: the oracle (one
    thread) on the same sample."""
    nz, nx, h = 501, 501, 10.0
    zc = (np.arange(nz - 1) + 0.5) * h
    xc = (np.arange(nx - 1) + 0.5) * h
    pert = 1.0 + 0.05 * np.sin(xc[:, None] / 700.0) * np.cos(zc[None, :] / 450.0)
    vel = np.ascontiguousarray(((1500.0 + 0.8 * zc)[None, :] * pert).ravel())
    rng = np.random.default_rng(21)
    zs = rng.uniform(0.0, (nz - 2) * h, nsrc)
    xs = rng.uniform(0.0, (nx - 2) * h, nsrc)
    zr = np.zeros(32)
    xr = np.linspace(0.0, (nx - 1) * h, 32)
    dp = ctypes.POINTER(c_double)
    ierr = c_int(1)
    L.fteik_solver2d_initialize(nz, nx, 0.0, 0.0, h, h, 2, 3.0, 0.0, 0,
                                byref(ierr))
    assert ierr.value == 0, "2D initialize"
    L.fteik_solver2d_setCellVelocityModel64f(vel.size, 0, vel.ctypes.data_as(
        dp), byref(ierr))
    L.fteik_solver2d_setSources64f(nsrc, zs.ctypes.data_as(dp),
                                   xs.ctypes.data_as(dp), byref(ierr))
    L.fteik_solver2d_setReceivers64f(zr.size, zr.ctypes.data_as(dp),
                                     xr.ctypes.data_as(dp), byref(ierr))
    assert ierr.value == 0, "2D set-up"
    L.fteik_solver2d_solveLSM(byref(ierr))          # warm-up
    assert ierr.value == 0, "2D solve"
    reps = 3
    hip_sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        L.fteik_solver2d_solveLSM(byref(ierr))
    hip_sync()
    sec = (time.perf_counter() - t0) / reps
    assert ierr.value == 0, "2D solve"
    ttr = np.zeros(zr.size * nsrc)
    L.fteik_solver2d_getTravelTimes64f(zr.size, ttr.ctypes.data_as(dp),
                                       byref(ierr))
    ttr = ttr.reshape(nsrc, zr.size)
    # nominal updates: initialization plus nsweep iterations, 4 sweeps over
    # the (nz-1)(nx-1) updated nodes each
    upd = nsrc * 3 * 4 * (nz - 1) * (nx - 1)
    out = {"workload": "2D, 501x501 h=10 m, v=1500+0.8z +-5% smooth, 256 "
                       "sources (one launch), nsweep=2, eps=3",
           "sources": nsrc, "seconds": sec, "value": nsrc / sec,
           "unit": "sources/s", "nominal_Gupdates_per_s": upd / sec / 1e9}
    from oracle import oracle as orc
    slow = orc.cell_slowness(vel, nz, nx)
    grid = (nz, nx, 1, h, h, 1.0, 0.0, 0.0, 0.0)
    same = True
    t0 = time.perf_counter()
    for k in range(cpu_sources):
        src = orc.source_locate(grid, zs[k], xs[k], slow=slow, is3d=False)
        ref = orc.solve2d(slow, nz, nx, h, h, src, 2, 3.0)
        same = same and bool(np.array_equal(
            ttr[k], orc.receivers(ref, grid, zr, xr, is3d=False)))
    tcpu = time.perf_counter() - t0
    port = cpu_sources / tcpu
    refr = ref_solver2d_rate(nz, nx, h, vel, zs[:cpu_sources],
                             xs[:cpu_sources], 2, 3.0)
    out["cpu_baseline"] = {"value": refr if refr else port,
                           "unit": "sources/s", "cores": 1,
                           "kind": "reference" if refr else "port",
                           "cpu_model": cpu_model(),
                           "sample": "%d sources" % cpu_sources,
                           "port_value": port,
                           "reference_over_port": (refr / port
                                                   if refr else None)}
    out["gpu_equals_oracle_on_sample"] = same
    return out


def c1_section(L, args, reps=20):
    """The reference's CPU-sized case: 101^3, h = 10 m, v(z) = 3000 + 0.5 z,
    source at the centre, 20 seeded surface receivers, 2 iterations, FP64
    dense (every update performed).  Parity: the field equals the oracle's
    bit for bit; CPU baseline: the oracle's OpenMP solve of the same case."""
    n, h, nsweep = 101, 10.0, 2
    zc = (np.arange(n - 1) + 0.5) * h
    vel = np.ascontiguousarray(np.tile(3000.0 + 0.5 * zc, (n - 1) * (n - 1)))
    zs = xs = ys = 50 * h
    rng = np.random.default_rng(101)
    xr = rng.uniform(0.0, (n - 1) * h, 20)
    yr = rng.uniform(0.0, (n - 1) * h, 20)
    zr = np.zeros(20)
    dp = ctypes.POINTER(c_double)
    ierr = c_int(1)
    L.fteik_solver3d_initialize(n, n, n, 0.0, 0.0, 0.0, h, h, h, nsweep, 5.0,
                                0.0, 0, byref(ierr))
    assert ierr.value == 0, "C1 initialize"
    L.fteik_solver3d_setTileSkipping(0, byref(ierr))
    L.fteik_solver3d_setVelocityModel64f(vel.size, vel.ctypes.data_as(dp),
                                         byref(ierr))
    z, x, y = (np.array([v]) for v in (zs, xs, ys))
    L.fteik_solver3d_setSources64f(1, z.ctypes.data_as(dp),
                                   x.ctypes.data_as(dp), y.ctypes.data_as(dp),
                                   byref(ierr))
    L.fteik_solver3d_setReceivers64f(20, zr.ctypes.data_as(dp),
                                     xr.ctypes.data_as(dp),
                                     yr.ctypes.data_as(dp), byref(ierr))
    assert ierr.value == 0, "C1 set-up"
    L.fteik_solver3d_solveSourceLSM(1, byref(ierr))          # warm-up
    hip_sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        L.fteik_solver3d_solveSourceLSM(1, byref(ierr))
    hip_sync()
    sec = (time.perf_counter() - t0) / reps
    assert ierr.value == 0, "C1 solve"
    ms, upd, corr = c_double(), c_double(), c_int()
    L.fteik_solver3d_getLastSolveInfo(byref(ms), byref(corr), byref(upd),
                                      byref(ierr))
    tt = np.zeros(n ** 3)
    L.fteik_solver3d_getTravelTimeField64f(tt.size, 0, tt.ctypes.data_as(dp),
                                           byref(ierr))
    out = {"workload": "101^3, h=10 m, v=3000+0.5z, source at the centre, "
                       "20 surface receivers, nsweep=2, FP64 dense",
           "ms_per_solve": sec * 1e3, "value": upd.value / sec / 1e9,
           "unit": "Gupdates/s", "updates_per_solve": upd.value}
    if not args.no_cpu:
        from oracle import oracle as orc
        slow = orc.cell_slowness(vel, n, n, n)
        src = orc.source_locate((n, n, n, h, h, h, 0.0, 0.0, 0.0), zs, xs, ys,
                                slow)
        t0 = time.perf_counter()
        ref = orc.solve3d(slow, n, n, n, h, h, h, src, nsweep, 5.0,
                          schedule=1)
        tcpu = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": upd.value / tcpu / 1e9,
                               "unit": "Gupdates/s", "cores": cpu_cores(),
                               "kind": "port", "cpu_model": cpu_model(),
                               "sample": "one complete solve", "seconds": tcpu}
        out["gpu_equals_oracle"] = bool(np.array_equal(ref, tt))
    L.fteik_solver3d_free()
    return out


def locator_section(L, args, nev=10000, cpu_events=16):
    """Full grid search over 48 FP32 station tables built by the batched
    solver and handed to the locator on the device."""
    (nz, nx, ny, h), vel, stations = locator_workload()
    ngrd = nz * nx * ny
    ierr = c_int(1)
    dp = ctypes.POINTER(c_double)
    ip = ctypes.POINTER(c_int)
    nsta = len(stations)
    # station tables: the reciprocal station-as-source driver, batched on
    # the 3D solver's grid and model, handed to the locator on the device
    L.fteik_solver3d_initialize(nz, nx, ny, 0.0, 0.0, 0.0, h, h, h, 2, 5.0,
                                0.0, 0, byref(ierr))
    assert ierr.value == 0
    L.fteik_solver3d_setVelocityModel64f(vel.size, vel.ctypes.data_as(dp),
                                         byref(ierr))
    assert ierr.value == 0
    L.locate_initialize(nev, nsta, ngrd, byref(ierr))
    assert ierr.value == 0
    L.locate_setReferenceCompatibility(0, byref(ierr))
    sz, sx, sy = (np.ascontiguousarray([s[i] for s in stations])
                  for i in range(3))
    L.locate_setTravelTimeFieldsFromStations3d(
        nsta, sz.ctypes.data_as(dp), sx.ctypes.data_as(dp),
        sy.ctypes.data_as(dp), 2, 5.0, 0, 1, byref(ierr))
    assert ierr.value == 0, "station tables"
    ms_tab, _, _ = batch_info(L)
    tables = np.zeros((nsta, ngrd))
    for k in range(nsta):
        tt = np.zeros(ngrd)
        L.fteik_batch3d_getTravelTimeField64f(k + 1, ngrd, 0,
                                              tt.ctypes.data_as(dp),
                                              byref(ierr))
        tables[k] = tt
    L.fteik_batch3d_free()
    L.fteik_solver3d_free()
    hypo = np.random.default_rng(10).integers(0, ngrd, nev)
    t0 = np.random.default_rng(11).uniform(0.0, 100.0, nev)
    noise = np.random.default_rng(12).normal(0.0, 0.05, (nev, nsta))
    picks = tables[:, hypo].T + t0[:, None] + noise
    o2t = np.arange(1, nsta + 1, dtype=np.int32)
    w = np.ones(nsta)
    for e in range(nev):
        p = np.ascontiguousarray(picks[e])
        L.locate_setObservation64f(e + 1, nsta, False, 0.0,
                                   o2t.ctypes.data_as(ip),
                                   p.ctypes.data_as(dp), w.ctypes.data_as(dp),
                                   byref(ierr))
    ev = np.arange(1, nev + 1, dtype=np.int32)
    opt = np.zeros(nev, dtype=np.int32)
    t0o = np.zeros(nev)
    obj = np.zeros(nev)
    call = (nev, ev.ctypes.data_as(ip), opt.ctypes.data_as(ip),
            t0o.ctypes.data_as(dp), obj.ctypes.data_as(dp), byref(ierr))
    L.locate_locateEvents(*call)               # warm-up
    hip_sync()
    tstart = time.perf_counter()
    L.locate_locateEvents(*call)
    hip_sync()
    tgpu = time.perf_counter() - tstart
    assert ierr.value == 0
    flops = 9.0 * nsta * ngrd * nev
    out = {"workload": "48 stations, 161x161x81 h=250 m, v=4000+0.4z, "
                       "10,000 events, sigma=0.05 s, full grid search",
           "events": nev, "seconds": tgpu, "value": nev / tgpu,
           "unit": "events/s", "tables_ms": ms_tab,
           "achieved_TFLOPs": flops / tgpu / 1e12, "fp32_peak_TFLOPs": 157.3,
           # the search issues unfused, unpacked FP32 operations (bitwise
           # parity): a wave64 FP32 instruction issues in 2 cycles (32 lanes
           # per SIMD per cycle, the rate of the FP32 peak counted as FMAs),
           # 4 SIMDs per CU, 256 CUs at 2.4 GHz
           "valu_issue_peak_Tops": 78.6,
           "valu_issue_frac": flops / tgpu / 78.6e12,
           "argmin_equals_true_cell": float(np.mean(
               np.abs(opt - 1 - hypo) == 0))}
    # CPU baseline and parity on a sample of events (oracle, fresh state)
    from oracle import oracle as orc
    cores = cpu_cores()
    loc = orc.Locator(nev, nsta, ngrd)
    for k in range(nsta):
        loc.set_table(k + 1, tables[k])
    same = 0
    tcpu = 0.0
    for e in range(cpu_events):
        loc.set_observation(e + 1, o2t, picks[e], w)
        tstart = time.perf_counter()
        r = loc.locate(e + 1, accumulate_t0=False, pad_argmin=False)
        tcpu += time.perf_counter() - tstart
        same += r == (int(opt[e]), float(t0o[e]), float(obj[e]))
    port = cpu_events / tcpu
    refr = ref_locator_rate(tables, picks, cpu_events, cores)
    out["cpu_baseline"] = {"value": refr if refr else port,
                           "unit": "events/s", "cores": cores,
                           "kind": "reference" if refr else "port",
                           "cpu_model": cpu_model(),
                           "sample": "%d events" % cpu_events,
                           "port_value": port,
                           "reference_over_port": (refr / port
                                                   if refr else None)}
    out["gpu_equals_oracle_on_sample"] = same == cpu_events
    out["surrogate"] = surrogate_section(L, stations, (nz, nx, ny, h), nev,
                                         args, opt, tgpu)
    L.locate_setReferenceCompatibility(1, byref(ierr))
    L.locate_finalize()
    return out


# parameters of the 3 -> 256 -> 256 -> 256 -> 256 -> 48 surrogate MLP
SUR_PARAMS = (3 * 256 + 256) + 3 * (256 * 256 + 256) + (256 * 48 + 48)


def surrogate_section(L, stations, grid, nev, args, full_opt, full_s):
    """Learned locator: MLP 3 -> 256^4 -> 48 trained on 2M (cell, times)
    pairs from the locator's tables (seed 13), then Gauss-Newton proposals
    scored with the exact tables; agreement and events/s against the full
    search."""
    nz, nx, ny, h = grid
    ierr = c_int(1)
    dp = ctypes.POINTER(c_double)
    ip = ctypes.POINTER(c_int)
    nsta = len(stations)
    sz = np.array([s[0] for s in stations])
    sx = np.array([s[1] for s in stations])
    sy = np.array([s[2] for s in stations])
    L.fteik_surrogate_initialize(nsta, nz, nx, ny, h, h, h,
                                 sz.ctypes.data_as(dp), sx.ctypes.data_as(dp),
                                 sy.ctypes.data_as(dp), 6000.0, 1.0, 13,
                                 byref(ierr))
    assert ierr.value == 0
    loss = c_double()
    hip_sync()
    t0 = time.perf_counter()
    L.fteik_surrogate_train(args.sur_steps, 8192, 2e-3, 2000000, byref(loss),
                            byref(ierr))
    hip_sync()
    ttrain = time.perf_counter() - t0
    assert ierr.value == 0
    ev = np.arange(1, nev + 1, dtype=np.int32)
    opt = np.zeros(nev, dtype=np.int32)
    t0o = np.zeros(nev)
    obj = np.zeros(nev)
    nfb = c_int()
    call = (nev, ev.ctypes.data_as(ip), 8, 6, opt.ctypes.data_as(ip),
            t0o.ctypes.data_as(dp), obj.ctypes.data_as(dp), byref(nfb),
            byref(ierr))
    L.locate_locateEventsSurrogate(*call)            # warm-up
    hip_sync()
    t0 = time.perf_counter()
    L.locate_locateEventsSurrogate(*call)
    hip_sync()
    tsur = time.perf_counter() - t0
    assert ierr.value == 0
    L.fteik_surrogate_free()
    return {"workload": "MLP 3-256-256-256-256-48 SiLU, BF16 MFMA training "
                        "on 2M table samples, 8 Gauss-Newton starts x 6 "
                        "steps, exact 4x4x4 cube scoring",
            "train_steps": args.sur_steps, "train_batch": 8192,
            "train_seconds": ttrain, "final_loss": loss.value,
            # 2 (forward) + 4 (backward) flops per parameter and row
            "train_TFLOPs": 6.0 * SUR_PARAMS * 8192 * args.sur_steps
                            / ttrain / 1e12,
            "value": nev / tsur, "unit": "events/s", "seconds": tsur,
            "speedup_vs_full_search": full_s / tsur,
            "agreement_with_full_search": float(np.mean(opt == full_opt)),
            "fallback_events": nfb.value}


def ref_locator_rate(tables, picks, nsample, cores):
    """events/s of the reference's own OpenMP grid search
    (locate.F90:112-126, :620-744) built from its sources (oracle/_ref),
    on the first nsample events; None when the build is absent."""
    from oracle import oracle as orc
    if not orc.have_ref():
        return None
    from ctypes import c_bool
    R = orc.ref()
    dp = ctypes.POINTER(c_double)
    ip = ctypes.POINTER(c_int)
    nsta, ngrd = tables.shape
    ierr = c_int(1)
    R.locate_initialize(c_int(nsample), c_int(nsta), c_int(ngrd), byref(ierr))
    if ierr.value != 0:
        return None
    for k in range(nsta):
        t = np.ascontiguousarray(tables[k])
        R.locate_setTravelTimeField64f(c_int(ngrd), c_int(k + 1),
                                       t.ctypes.data_as(dp), byref(ierr))
    o2t = np.arange(1, nsta + 1, dtype=np.int32)
    w = np.ones(nsta)
    for e in range(nsample):
        p = np.ascontiguousarray(picks[e])
        R.locate_setObservation64f(c_int(e + 1), c_int(nsta), c_bool(False),
                                   c_double(0.0), o2t.ctypes.data_as(ip),
                                   p.ctypes.data_as(dp), w.ctypes.data_as(dp),
                                   byref(ierr))
    opt, t0, ob = c_int(), c_double(), c_double()
    tstart = time.perf_counter()
    for e in range(nsample):
        R.locate_locateEvent(c_int(e + 1), byref(opt), byref(t0), byref(ob))
    sec = time.perf_counter() - tstart
    R.locate_finalize()
    return nsample / sec


def ref_solver2d_rate(nz, nx, h, vel, zs, xs, nsweep, eps):
    """sources/s of the reference's own 2D level-set solver (single-
    threaded, solver2D.F90:782-978) built from its sources (oracle/_ref);
    None when the build is absent."""
    from oracle import oracle as orc
    if not orc.have_ref():
        return None
    R = orc.ref()
    dp = ctypes.POINTER(c_double)
    ierr = c_int(1)
    R.fteik_solver2d_initialize(c_int(nz), c_int(nx), c_double(0.0),
                                c_double(0.0), c_double(h), c_double(h),
                                c_int(nsweep), c_double(eps), c_double(0.0),
                                c_int(0), byref(ierr))
    if ierr.value != 0:
        return None
    v = np.ascontiguousarray(vel)
    R.fteik_solver2d_setCellVelocityModel64f(c_int(v.size), c_int(0),
                                             v.ctypes.data_as(dp), byref(ierr))
    z, x = np.ascontiguousarray(zs), np.ascontiguousarray(xs)
    R.fteik_solver2d_setSources64f(c_int(z.size), z.ctypes.data_as(dp),
                                   x.ctypes.data_as(dp), byref(ierr))
    tstart = time.perf_counter()
    for k in range(z.size):
        R.fteik_solver2d_solveSourceLSM(c_int(k + 1), byref(ierr))
    sec = time.perf_counter() - tstart
    R.fteik_solver2d_free()
    return z.size / sec if ierr.value == 0 else None


def cpu_model():
    """The host CPU's model name (lscpu), for the baseline records."""
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True,
                             timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return "unknown"


_CORES = None


def cpu_cores():
    """Host threads for the CPU baselines: the cores this process may run
    on at start-up, at most 16 (the host share of one GPU on the benchmark
    machines).  Taken once, before any OpenMP runtime starts: a runtime
    that binds its threads narrows the main thread's affinity, which a
    later query would report."""
    global _CORES
    if _CORES is None:
        try:
            n = len(os.sched_getaffinity(0))
        except AttributeError:
            n = os.cpu_count() or 1
        _CORES = max(1, min(16, n))
    return _CORES


def hip_sync():
    hip = ctypes.CDLL("libamdhip64.so.7")
    return hip.hipDeviceSynchronize()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the batched-solve and locator sections")
    ap.add_argument("--sur-steps", type=int, default=3000,
                    help="training steps of the locator surrogate")
    ap.add_argument("--only-locator", action="store_true",
                    help="run only the locator and surrogate sections "
                         "(for profiling)")
    args = ap.parse_args()
    # one thread count for every OpenMP runtime the baselines load (the
    # oracle's and the reference build's), fixed before either starts; no
    # thread binding, so the reference's runtime, which starts later, still
    # sees every core
    os.environ["OMP_NUM_THREADS"] = str(cpu_cores())
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not "
              "supported (the solver has no multi-GPU path)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import fteik_amd
    L = fteik_amd.load_library()
    if args.only_locator:
        print(json.dumps({"locator": locator_section(L, args)}), flush=True)
        return
    vel, (zs, xs, ys) = layered_model()
    ierr = c_int(1)

    def init(precision, nsweep, skipping=0, overlap=0):
        L.fteik_solver3d_initialize(N, N, N, 0.0, 0.0, 0.0, H, H, H, nsweep,
                                    5.0, 0.0, 0, byref(ierr))
        assert ierr.value == 0, "initialize"
        L.fteik_solver3d_setTileSkipping(skipping, byref(ierr))
        L.fteik_solver3d_setSweepOverlap(overlap, byref(ierr))
        assert ierr.value == 0, "sweep options"
        L.fteik_solver3d_setVelocityModel64f(vel.size, vel.ctypes.data_as(
            ctypes.POINTER(c_double)), byref(ierr))
        assert ierr.value == 0, "model"
        z, x, y = (np.array([v]) for v in (zs, xs, ys))
        dp = ctypes.POINTER(c_double)
        L.fteik_solver3d_setSources64f(1, z.ctypes.data_as(dp),
                                       x.ctypes.data_as(dp),
                                       y.ctypes.data_as(dp), byref(ierr))
        assert ierr.value == 0, "sources"
        L.fteik_solver3d_setPrecision(precision, byref(ierr))
        assert ierr.value == 0, "precision"

    def run(k):
        for _ in range(k):
            L.fteik_solver3d_solveSourceLSM(1, byref(ierr))
            if ierr.value != 0:
                raise RuntimeError("solve failed")

    def info():
        ms, upd, corr = c_double(), c_double(), c_int()
        e = c_int()
        L.fteik_solver3d_getLastSolveInfo(byref(ms), byref(corr), byref(upd),
                                          byref(e))
        return ms.value, corr.value, upd.value

    def field():
        tt = np.zeros(N ** 3)
        L.fteik_solver3d_getTravelTimeField64f(tt.size, 0, tt.ctypes.data_as(
            ctypes.POINTER(c_double)), byref(ierr))
        assert ierr.value == 0
        return tt

    def sweep_ms(precision, nrep=4):
        ms = c_double()
        L.fteik_amd_timeSweeps3d(precision, nrep, byref(ms), byref(ierr))
        assert ierr.value == 0, "timeSweeps3d"
        return ms.value

    def timed(k, w):
        run(w)
        hip_sync()
        t0 = time.perf_counter()
        run(k)
        hip_sync()
        return (time.perf_counter() - t0) / k

    # ------------------------------------------- FP64 iterations to converge
    # The reference runs a fixed number of iterations (no convergence test);
    # the workload uses the smallest count after which one more iteration
    # changes no node.
    prev, nsweep = None, None
    for ns in range(1, 41):
        init(0, ns)
        run(1)
        cur = field()
        if prev is not None and np.array_equal(cur, prev):
            nsweep = ns - 1
            break
        prev = cur
    if nsweep is None:
        nsweep = 40
    print("bench: FP64 converges after %d iterations" % nsweep,
          file=sys.stderr, flush=True)

    # ---------------------------------------------------------- FP64 path
    # Headline: every update of every sweep performed (no tile skipping).
    init(0, nsweep)
    t64 = timed(args.steps, args.warmup)
    _, _, upd64 = info()
    tt64 = field()
    print("bench: fp64 solve %.3f ms" % (t64 * 1e3), file=sys.stderr,
          flush=True)
    ms_launch64 = sweep_ms(0)

    # Same solve with exact tile skipping (the library default) and with
    # chained sweeps: time to the same field, not an update rate.
    init(0, nsweep, skipping=1)
    tskip = timed(args.steps, args.warmup)
    skip_equal = bool(np.array_equal(field(), tt64))
    init(0, nsweep, overlap=2)
    tchain = timed(args.steps, args.warmup)
    chain_equal = bool(np.array_equal(field(), tt64))
    print("bench: fp64 with tile skipping %.3f ms (equal %s), chained "
          "sweeps %.3f ms (equal %s)" % (tskip * 1e3, skip_equal,
                                         tchain * 1e3, chain_equal),
          file=sys.stderr, flush=True)

    # --------------------------------------------------------- mixed path
    # library defaults for the FP64 corrections (tile skipping on)
    init(1, nsweep, skipping=1)
    tmix = timed(args.steps, args.warmup)
    _, corr, updmix = info()
    ttm = field()
    rel = np.abs(ttm - tt64) / np.maximum(tt64, 1e-12)
    print("bench: mixed solve %.3f ms, %d FP64 correction iterations, max "
          "rel err %.3g" % (tmix * 1e3, corr, rel.max()), file=sys.stderr,
          flush=True)
    ms_launch32 = sweep_ms(1)

    # latency floor of a sweep: one level after another, each at least one
    # dependent iteration stencil (one wave per SIMD)
    us_chain = c_double()
    L.fteik_amd_timeStencilChain3d(0, 4000, byref(us_chain), byref(ierr))
    assert ierr.value == 0, "stencil chain"
    levels = 3 * (N - 1) - 2             # flipped i+j+k over the sweep box
    floor_ms = levels * us_chain.value * 1e-3

    units = float((N - 1) ** 3)          # updates per Gauss-Seidel launch
    t1, t2, nt1 = c_int(), c_int(), c_int()
    L.fteik_solver3d_getTile(byref(t1), byref(t2), byref(nt1), byref(ierr))
    tile64 = t1.value
    ach64 = BYTES_PER_UPDATE["f64"] * units / (ms_launch64 * 1e-3) / 1e9
    traffic = None
    if os.path.isfile(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("fp64_bytes_per_launch")

    # ------------------------------------------------------- CPU baseline
    cpu = None
    if not args.no_cpu:
        cores = cpu_cores()
        from oracle import oracle as orc
        slow = orc.cell_slowness(vel, N, N, N)
        src = orc.source_locate((N, N, N, H, H, H, 0.0, 0.0, 0.0), zs, xs, ys,
                                slow)
        t0 = time.perf_counter()
        ref = orc.solve3d(slow, N, N, N, H, H, H, src, nsweep, 5.0,
                          schedule=1)
        tcpu = time.perf_counter() - t0
        cpu = {"value": upd64 / tcpu / 1e9, "unit": "Gupdates/s",
               "cores": cores, "kind": "port", "cpu_model": cpu_model(),
               "sample": "one complete FP64 solve of the same 257^3 "
                         "workload (oracle, OpenMP level-synchronous)",
               "seconds": tcpu,
               "fp64_gpu_equals_oracle": bool(np.array_equal(ref, tt64))}

    out = {
        "metric": "Gupdates/s at 257^3 FP64 vs FP32-mixed",
        "value": upd64 / t64 / 1e9,
        "unit": "Gupdates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t64 * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "3D single-source level-set sweeps, 257^3, "
                               "h=20 m, 8 layers 1500-6500 m/s +-5% smoothed "
                               "perturbation, FP64 iterations to "
                               "convergence",
                   "grid": [N, N, N], "h": H, "nsweep": nsweep, "eps": 5.0,
                   "source_zxy": [zs, xs, ys],
                   "updates_per_solve": upd64},
        "fp64_tile_skipping": {
            "ms_per_solve": tskip * 1e3, "speedup_vs_dense": t64 / tskip,
            "field_equals_dense": skip_equal,
            "note": "same solve, tiles whose pass would reproduce their "
                    "values exactly are skipped; time to the same field"},
        "fp64_chained_sweeps": {
            "ms_per_solve": tchain * 1e3, "speedup_vs_dense": t64 / tchain,
            "field_equals_dense": chain_equal,
            "note": "all 8*nsweep iteration sweeps in one launch, each "
                    "overlapping the one before it; every update performed"},
        "mixed": {"value": upd64 / tmix / 1e9, "unit": "Gupdates/s",
                  "ms_per_step": tmix * 1e3, "speedup_vs_fp64": t64 / tmix,
                  "fp32_iterations": nsweep,
                  "speedup_vs_fp64_tile_skipping": tskip / tmix,
                  "fp64_correction_iterations": corr,
                  "updates_performed": updmix,
                  "max_rel_err_vs_fp64": float(rel.max()),
                  "median_rel_err_vs_fp64": float(np.median(rel)),
                  "fp32_sweep_ms_per_launch": ms_launch32},
        "roofline": {"kernel": "FP64 Gauss-Seidel sweep, tile %d" % tile64,
                     "bound": "hbm", "achieved": ach64, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": ach64 / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "ms_per_launch": ms_launch64,
                     "bytes_per_update": BYTES_PER_UPDATE["f64"],
                     "updates_per_launch": units,
                     "critical_path": {
                         "levels_per_launch": levels,
                         "stencil_chain_us_per_level": us_chain.value,
                         "floor_ms_per_launch": floor_ms,
                         "frac": floor_ms / ms_launch64,
                         "note": "a sweep's levels are sequential; each "
                                 "costs at least one dependent FP64 "
                                 "iteration stencil (measured live, one "
                                 "wave per SIMD)"}},
        "cpu_baseline": cpu,
    }
    if not args.no_extra:
        out["c1_101"] = c1_section(L, args)
        print("bench: 101^3 %.3f ms/solve" % out["c1_101"]["ms_per_solve"],
              file=sys.stderr, flush=True)
        out["batch"] = batch_section(L, args)
        print("bench: batch %.1f Gupd/s" % out["batch"]["value"],
              file=sys.stderr, flush=True)
        out["locator"] = locator_section(L, args)
        print("bench: locator %.0f events/s" % out["locator"]["value"],
              file=sys.stderr, flush=True)
        out["solver2d"] = solver2d_section(L)
        print("bench: 2D %.0f sources/s" % out["solver2d"]["value"],
              file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
