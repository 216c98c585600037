"""GPU parity at the benchmark workloads' own sizes (the 101^3 linear
gradient solve is in test_gpu_solver3d.py):

- 257^3 layered model, h = 20 m: the FP64 field equals the oracle's bit for
  bit after the iteration count at which one more iteration changes nothing;
  the mixed mode (FP32 iterations, FP64 corrections) is within 1e-6
  relative of it;
- 64 grids of 64..192 per axis in one batch: a seeded sample of 8 grids
  equals the oracle bit for bit in FP64, and the FP32-mixed batch is within
  1e-6 relative of the converged FP64 batch on every grid;
- 48 station tables of 161 x 161 x 81 from the batched solver, 10,000
  events: the full grid search equals the oracle's on a seeded sample of
  240 events bit for bit;
- the surrogate-pruned search on the same 10,000 events agrees with the full
  search on >= 99.5 % of them at >= 10x its events/s, and against the same
  search seeded from straight-ray proposals alone (no network) it gains in
  agreement or in events/s and loses in neither.

All calls go through the C-ABI; the workloads are the benchmark's own
generators (bench.py), seeded.
"""
import time
from ctypes import byref, c_double, c_int

import numpy as np
import pytest

from oracle import oracle as orc
from tests.gpu_helpers import GPU3D, lib, oracle3d, _d, _i

pytestmark = pytest.mark.gpu


def _say(*a):
    print("case:", *a, flush=True)


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-12)))


# ------------------------------------------------------------------ 257^3
def test_layered_257_fp64_bitwise_and_mixed():
    from bench import layered_model, N, H
    nsweep = 19
    vel, (zs, xs, ys) = layered_model()
    _say("257^3 layered FP64, nsweep", nsweep)
    g = GPU3D(N, N, N, H, nsweep=nsweep, eps=5.0)
    g.set_cells(vel)
    g.set_sources(zs, xs, ys)
    assert g.solve(1) == 0
    t64 = g.field()
    L = lib()
    ierr = c_int(1)
    L.fteik_solver3d_setNumberOfSweeps(nsweep + 1, byref(ierr))
    assert g.solve(1) == 0
    # converged: one more iteration changes no node
    np.testing.assert_array_equal(g.field(), t64)
    _say("257^3 oracle")
    ref, _, _ = oracle3d(vel, N, N, N, H, zs, xs, ys, nsweep, 5.0)
    np.testing.assert_array_equal(t64, ref)
    del ref
    _say("257^3 mixed")
    gm = GPU3D(N, N, N, H, nsweep=nsweep, eps=5.0, precision=1)
    gm.set_cells(vel)
    gm.set_sources(zs, xs, ys)
    assert gm.solve(1) == 0
    err = _rel(gm.field(), t64)
    print("257^3 mixed: max rel err vs FP64 %.3g, corrections %d"
          % (err, gm.info()[1]), flush=True)
    assert err <= 1e-6


# ------------------------------------------------------------ 64-grid batch
def _batch(L, models, h, nsweep, precision, maxcorr=64):
    ierr = c_int(1)
    L.fteik_batch3d_initialize(len(models), byref(ierr))
    assert ierr.value == 0
    for k, ((nz, nx, ny), vel, (zs, xs, ys)) in enumerate(models):
        L.fteik_batch3d_setGrid(k + 1, nz, nx, ny, 0.0, 0.0, 0.0, h, h, h,
                                _d(vel), zs, xs, ys, byref(ierr))
        assert ierr.value == 0, k
    L.fteik_batch3d_setPrecision(precision, byref(ierr))
    L.fteik_batch3d_setMixedParameters(1.0 / 64.0, maxcorr, byref(ierr))
    assert ierr.value == 0
    L.fteik_batch3d_solve(nsweep, 5.0, byref(ierr))
    assert ierr.value == 0
    out = []
    for k, ((nz, nx, ny), _, _) in enumerate(models):
        tt = np.zeros(nz * nx * ny)
        L.fteik_batch3d_getTravelTimeField64f(k + 1, tt.size, 0, _d(tt),
                                              byref(ierr))
        assert ierr.value == 0
        out.append(tt)
    ms, upd, corr = c_double(), c_double(), c_int()
    L.fteik_batch3d_getLastSolveInfo(byref(ms), byref(corr), byref(upd),
                                     byref(ierr))
    L.fteik_batch3d_free()
    return out, corr.value


def test_batch_64_grids_fp64_sample_and_mixed():
    from bench import batch_models
    h = 50.0
    models = batch_models()
    L = lib()
    _say("batch 64 grids FP64 nsweep 2")
    f2, _ = _batch(L, models, h, 2, 0)
    sample = np.sort(np.random.default_rng(55).choice(64, 8, replace=False))
    for k in sample:
        (nz, nx, ny), vel, (zs, xs, ys) = models[k]
        _say("batch grid", int(k), (nz, nx, ny))
        ref, _, _ = oracle3d(vel, nz, nx, ny, h, zs, xs, ys, 2, 5.0)
        np.testing.assert_array_equal(f2[k], ref)
    del f2
    _say("batch 64 grids FP64 converged")
    conv, _ = _batch(L, models, h, 16, 0)
    more, _ = _batch(L, models, h, 17, 0)
    for a, b in zip(conv, more):
        np.testing.assert_array_equal(a, b)
    del more
    _say("batch 64 grids FP32-mixed")
    mixed, corr = _batch(L, models, h, 8, 1)
    worst = max(_rel(m, c) for m, c in zip(mixed, conv))
    print("batch mixed: %d FP64 corrections, max rel err %.3g"
          % (corr, worst), flush=True)
    assert worst <= 1e-6


# ---------------------------------------------------- locator and surrogate
NEV = 10000


@pytest.fixture(scope="module")
def located():
    """48 station tables from the batched solver, handed to the locator on
    the device, 10,000 events set, full grid search run."""
    from bench import locator_workload
    (nz, nx, ny, h), vel, stations = locator_workload()
    ngrd = nz * nx * ny
    nsta = len(stations)
    L = lib()
    ierr = c_int(1)
    _say("station tables: batch of", nsta)
    L.fteik_batch3d_initialize(nsta, byref(ierr))
    for k, (zs, xs, ys) in enumerate(stations):
        L.fteik_batch3d_setGrid(k + 1, nz, nx, ny, 0.0, 0.0, 0.0, h, h, h,
                                _d(vel), zs, xs, ys, byref(ierr))
        assert ierr.value == 0
    L.fteik_batch3d_solve(2, 5.0, byref(ierr))
    assert ierr.value == 0
    L.locate_initialize(NEV, nsta, ngrd, byref(ierr))
    assert ierr.value == 0
    L.locate_setReferenceCompatibility(0, byref(ierr))
    tables = np.zeros((nsta, ngrd))
    for k in range(nsta):
        L.locate_setTravelTimeFieldFromBatch3d(k + 1, k + 1, byref(ierr))
        assert ierr.value == 0
        L.fteik_batch3d_getTravelTimeField64f(k + 1, ngrd, 0, _d(tables[k]),
                                              byref(ierr))
        assert ierr.value == 0
    L.fteik_batch3d_free()
    hypo = np.random.default_rng(10).integers(0, ngrd, NEV)
    t0 = np.random.default_rng(11).uniform(0.0, 100.0, NEV)
    noise = np.random.default_rng(12).normal(0.0, 0.05, (NEV, nsta))
    picks = tables[:, hypo].T + t0[:, None] + noise
    o2t = np.arange(1, nsta + 1, dtype=np.int32)
    w = np.ones(nsta)
    for e in range(NEV):
        p = np.ascontiguousarray(picks[e])
        L.locate_setObservation64f(e + 1, nsta, False, 0.0, _i(o2t), _d(p),
                                   _d(w), byref(ierr))
    ev = np.arange(1, NEV + 1, dtype=np.int32)
    full = (np.zeros(NEV, dtype=np.int32), np.zeros(NEV), np.zeros(NEV))
    call = (NEV, _i(ev), _i(full[0]), _d(full[1]), _d(full[2]), byref(ierr))
    L.locate_locateEvents(*call)                 # warm-up
    assert ierr.value == 0
    t = time.perf_counter()
    L.locate_locateEvents(*call)
    tfull = time.perf_counter() - t
    assert ierr.value == 0
    yield dict(L=L, grid=(nz, nx, ny, h), vel=vel, stations=stations,
               tables=tables, picks=picks, full=full, tfull=tfull, ev=ev)
    L.locate_setReferenceCompatibility(1, byref(ierr))
    L.locate_finalize()


def test_station_tables_and_full_search_match_oracle(located):
    d = located
    nz, nx, ny, h = d["grid"]
    ngrd = nz * nx * ny
    # two station tables against the oracle's single-grid solve
    for k in (0, 29):
        zs, xs, ys = d["stations"][k]
        _say("station table", k)
        ref, _, _ = oracle3d(d["vel"], nz, nx, ny, h, zs, xs, ys, 2, 5.0)
        np.testing.assert_array_equal(d["tables"][k], ref)
    nsta = len(d["stations"])
    loc = orc.Locator(NEV, nsta, ngrd)
    for k in range(nsta):
        loc.set_table(k + 1, d["tables"][k])
    o2t = np.arange(1, nsta + 1, dtype=np.int32)
    w = np.ones(nsta)
    sample = np.random.default_rng(77).choice(NEV, 240, replace=False)
    opt, t0, ob = d["full"]
    _say("full search vs oracle on", sample.size, "events")
    for e in sample:
        loc.set_observation(int(e) + 1, o2t, d["picks"][e], w)
        want = loc.locate(int(e) + 1, accumulate_t0=False, pad_argmin=False)
        assert (int(opt[e]), float(t0[e]), float(ob[e])) == want, int(e)


def _surrogate_run(L, ev, net):
    ierr = c_int(1)
    L.fteik_surrogate_setNetworkProposals(net, byref(ierr))
    assert ierr.value == 0
    n = ev.size
    out = (np.zeros(n, dtype=np.int32), np.zeros(n), np.zeros(n))
    nfb = c_int()
    call = (n, _i(ev), 8, 6, _i(out[0]), _d(out[1]), _d(out[2]), byref(nfb),
            byref(ierr))
    L.locate_locateEventsSurrogate(*call)        # warm-up
    assert ierr.value == 0
    t = time.perf_counter()
    L.locate_locateEventsSurrogate(*call)
    sec = time.perf_counter() - t
    assert ierr.value == 0
    return out, sec, nfb.value


def test_surrogate_agreement_speed_and_ablation(located):
    d = located
    L = d["L"]
    nz, nx, ny, h = d["grid"]
    st = d["stations"]
    nsta = len(st)
    sz = np.array([s[0] for s in st])
    sx = np.array([s[1] for s in st])
    sy = np.array([s[2] for s in st])
    ierr = c_int(1)
    L.fteik_surrogate_initialize(nsta, nz, nx, ny, h, h, h, _d(sz), _d(sx),
                                 _d(sy), 6000.0, 1.0, 13, byref(ierr))
    assert ierr.value == 0
    loss = c_double()
    _say("surrogate training")
    L.fteik_surrogate_train(3000, 8192, 2e-3, 2000000, byref(loss),
                            byref(ierr))
    assert ierr.value == 0
    full = d["full"]
    res = {}
    for net in (1, 0):
        _say("surrogate search, network", net)
        out, sec, nfb = _surrogate_run(L, d["ev"], net)
        agree = float(np.mean(out[0] == full[0]))
        same = out[0] == full[0]
        # where the cells agree, origin time and objective are the exact
        # tables' values
        np.testing.assert_array_equal(out[1][same], full[1][same])
        np.testing.assert_array_equal(out[2][same], full[2][same])
        res[net] = (agree, NEV / sec, nfb)
        print("network %d: agreement %.4f, %.0f events/s, %d fallbacks"
              % (net, agree, NEV / sec, nfb), flush=True)
    L.fteik_surrogate_setNetworkProposals(1, byref(ierr))
    L.fteik_surrogate_free()
    full_rate = NEV / d["tfull"]
    print("full search %.0f events/s" % full_rate, flush=True)
    agree, rate, _ = res[1]
    assert agree >= 0.995
    assert rate >= 10.0 * full_rate
    a0, r0, _ = res[0]
    assert agree >= a0 and rate >= r0 and (agree > a0 or rate > 1.05 * r0)
