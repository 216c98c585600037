"""GPU parity of the batched 3D solver: every grid of a batch equals the
oracle's single-grid solve bit for bit, for several tile shapes, and
locator tables filled from the batch on the device equal tables uploaded
from the host."""
from ctypes import byref, c_double, c_int

import numpy as np
import pytest

from oracle import oracle as orc
from tests.gpu_helpers import hetero, lib, oracle3d, _d, _i

pytestmark = pytest.mark.gpu

GRIDS = [
    # nz, nx, ny, h, source in cells
    (23, 19, 17, 15.0, (7.7, 10.2, 5.5)),
    (40, 37, 33, 12.0, (30.25, 3.5, 20.0)),
    (17, 50, 45, 10.0, (0.0, 47.2, 1.1)),
    (31, 18, 47, 20.0, (12.6, 15.9, 44.3)),
    (33, 33, 33, 10.0, (16.0, 16.0, 16.0)),
]


def _batch(grids, models, nsweep, eps, tile=None, precision=0):
    L = lib()
    ierr = c_int(1)
    L.fteik_batch3d_initialize(len(grids), byref(ierr))
    assert ierr.value == 0
    for g, ((nz, nx, ny, h, (zc, xc, yc)), vel) in enumerate(zip(grids,
                                                                 models)):
        v = np.ascontiguousarray(vel)
        L.fteik_batch3d_setGrid(g + 1, nz, nx, ny, 0.0, 0.0, 0.0, h, h, h,
                                _d(v), zc * h, xc * h, yc * h, byref(ierr))
        assert ierr.value == 0
    # an int, or a sequence of tiles set in turn (the last one counts)
    for t in ([] if tile is None else
              [tile] if isinstance(tile, int) else tile):
        L.fteik_batch3d_setTile(t, byref(ierr))
        assert ierr.value == 0
    L.fteik_batch3d_setPrecision(precision, byref(ierr))
    L.fteik_batch3d_solve(nsweep, eps, byref(ierr))
    assert ierr.value == 0
    return L


def _field(L, g, n):
    tt = np.zeros(n)
    ierr = c_int(1)
    L.fteik_batch3d_getTravelTimeField64f(g + 1, n, 0, _d(tt), byref(ierr))
    assert ierr.value == 0
    return tt


@pytest.mark.parametrize("tile", [0, 3, 6, 7, 8, 9, 10, 11, 12, 13, 18,
                                  (3, 12), (0, 10), (3, 18)])
def test_batch_matches_oracle(tile):
    print("case: batch tile", tile, flush=True)
    models = [hetero(nz, nx, ny, h, 40 + k)
              for k, (nz, nx, ny, h, _) in enumerate(GRIDS)]
    L = _batch(GRIDS, models, 2, 4.0, tile=tile)
    for k, ((nz, nx, ny, h, (zc, xc, yc)), vel) in enumerate(zip(GRIDS,
                                                                 models)):
        ref, _, _ = oracle3d(vel, nz, nx, ny, h, zc * h, xc * h, yc * h, 2,
                             4.0)
        np.testing.assert_array_equal(_field(L, k, nz * nx * ny), ref)
    ms, upd, corr, ierr = c_double(), c_double(), c_int(), c_int()
    L.fteik_batch3d_getLastSolveInfo(byref(ms), byref(corr), byref(upd),
                                     byref(ierr))
    assert ierr.value == 0 and upd.value > 0 and ms.value > 0
    L.fteik_batch3d_free()


@pytest.mark.parametrize("tile", [6, 13, 9, 18])
def test_large_batch_matches_oracle(tile):
    """48 grids of 65^3 (four sources, each twelve times): more workgroups
    per sweep than the GPU holds at once, so later grids' tiles wait for
    earlier ones to leave; every grid equals the oracle bit for bit."""
    print("case: large batch tile", tile, flush=True)
    n, h = 65, 25.0
    vel = hetero(n, n, n, h, 91)
    sources = [(3.3, 60.2, 31.0), (32.0, 32.0, 32.0), (61.5, 2.5, 7.25),
               (0.0, 44.4, 62.0)]
    grids = [(n, n, n, h, sources[k % 4]) for k in range(48)]
    L = _batch(grids, [vel] * 48, 2, 5.0, tile=tile)
    refs = [oracle3d(vel, n, n, n, h, zc * h, xc * h, yc * h, 2, 5.0)[0]
            for (zc, xc, yc) in sources]
    for k in range(48):
        np.testing.assert_array_equal(_field(L, k, n ** 3), refs[k % 4])
    L.fteik_batch3d_free()


def test_batch_tables_feed_locator():
    """Three station tables solved in one batch go to the locator on the
    device; the result equals tables uploaded from the host fields."""
    nz, nx, ny, h = 21, 24, 22, 100.0
    vel = hetero(nz, nx, ny, h, 77)
    stations = [(0.0, 3.3, 4.4), (0.0, 17.9, 2.1), (0.0, 9.5, 19.6)]
    grids = [(nz, nx, ny, h, s) for s in stations]
    L = _batch(grids, [vel] * 3, 2, 5.0)
    ngrd = nz * nx * ny
    fields = [_field(L, k, ngrd) for k in range(3)]
    ierr = c_int(1)
    rng = np.random.default_rng(5)
    nev = 6
    cells = rng.integers(0, ngrd, nev)
    picks = np.array([[f[c] + 10.0 for f in fields] for c in cells])
    o2t = np.arange(1, 4, dtype=np.int32)
    w = np.ones(3)
    results = []
    for source in ("batch", "host"):
        L.locate_initialize(nev, 3, ngrd, byref(ierr))
        assert ierr.value == 0
        L.locate_setReferenceCompatibility(0, byref(ierr))
        for k in range(3):
            if source == "batch":
                L.locate_setTravelTimeFieldFromBatch3d(k + 1, k + 1,
                                                       byref(ierr))
            else:
                L.locate_setTravelTimeField64f(ngrd, k + 1, _d(fields[k]),
                                               byref(ierr))
            assert ierr.value == 0
        for e in range(nev):
            p = np.ascontiguousarray(picks[e])
            L.locate_setObservation64f(e + 1, 3, False, 0.0, _i(o2t), _d(p),
                                       _d(w), byref(ierr))
        ev = np.arange(1, nev + 1, dtype=np.int32)
        opt = np.zeros(nev, dtype=np.int32)
        t0 = np.zeros(nev)
        ob = np.zeros(nev)
        L.locate_locateEvents(nev, _i(ev), _i(opt), _d(t0), _d(ob),
                              byref(ierr))
        assert ierr.value == 0
        results.append((opt.copy(), t0.copy(), ob.copy()))
        L.locate_setReferenceCompatibility(1, byref(ierr))
        L.locate_finalize()
    for a, b in zip(results[0], results[1]):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(results[0][0], cells + 1)
    L.fteik_batch3d_free()


def test_batch_error_paths():
    L = lib()
    ierr = c_int(0)
    L.fteik_batch3d_initialize(2, byref(ierr))
    assert ierr.value == 0
    v = np.full(9 * 9 * 9, 3000.0)
    L.fteik_batch3d_setGrid(3, 10, 10, 10, 0, 0, 0, 10, 10, 10, _d(v), 5, 5,
                            5, byref(ierr))
    assert ierr.value == 1                 # grid index out of range
    L.fteik_batch3d_setGrid(1, 10, 10, 10, 0, 0, 0, 10, 10, 10, _d(v), 95,
                            5, 5, byref(ierr))
    assert ierr.value == 1                 # source in the last cell
    L.fteik_batch3d_setGrid(1, 10, 10, 10, 0, 0, 0, 10, 10, 10, _d(v), 45,
                            45, 45, byref(ierr))
    assert ierr.value == 0
    L.fteik_batch3d_solve(2, 5.0, byref(ierr))
    assert ierr.value == 1                 # grid 2 never set
    L.fteik_batch3d_free()


@pytest.mark.parametrize("tile", [3, 6, 11, 13, 18])
def test_batch_mixed_precision_matches_fp64(tile):
    """Mixed mode (FP32 iterations, FP64 corrections to convergence) of a
    batch: every grid is within 1e-6
    relative of its converged FP64 batch result.  Not bitwise: the gated
    sweeps can settle at a few nodes on a slightly different fixed point
    from a different starting field (sweeps never raise a time)."""
    print("case: batch mixed tile", tile, flush=True)
    models = [hetero(nz, nx, ny, h, 60 + k)
              for k, (nz, nx, ny, h, _) in enumerate(GRIDS)]
    L = _batch(GRIDS, models, 30, 4.0, tile=6, precision=0)
    ref = [_field(L, k, nz * nx * ny)
           for k, (nz, nx, ny, h, _) in enumerate(GRIDS)]
    L.fteik_batch3d_free()
    L = _batch(GRIDS, models, 6, 4.0, tile=tile, precision=1)
    ms, upd, corr, ierr = c_double(), c_double(), c_int(), c_int()
    L.fteik_batch3d_getLastSolveInfo(byref(ms), byref(corr), byref(upd),
                                     byref(ierr))
    print("FP64 correction iterations", corr.value, flush=True)
    for k, (nz, nx, ny, h, _) in enumerate(GRIDS):
        got = _field(L, k, nz * nx * ny)
        rel = np.abs(got - ref[k]) / np.maximum(ref[k], 1e-12)
        print("grid", k, "differing nodes", int((got != ref[k]).sum()),
              "max rel", float(rel.max()), flush=True)
        assert rel.max() <= 1e-6
    L.fteik_batch3d_free()


def test_batch_autotune():
    """fteik_batch3d_autotune picks a tile by timing the batch; the batch is
    left solved and every grid equals the oracle."""
    models = [hetero(nz, nx, ny, h, 80 + k)
              for k, (nz, nx, ny, h, _) in enumerate(GRIDS)]
    L = _batch(GRIDS, models, 2, 4.0)
    t, ierr = c_int(-1), c_int(1)
    L.fteik_batch3d_autotune(2, 4.0, byref(t), byref(ierr))
    assert ierr.value == 0 and t.value >= 0
    print("case: batch autotuned tile", t.value, flush=True)
    for k, ((nz, nx, ny, h, (zc, xc, yc)), vel) in enumerate(zip(GRIDS,
                                                                 models)):
        ref, _, _ = oracle3d(vel, nz, nx, ny, h, zc * h, xc * h, yc * h, 2,
                             4.0)
        np.testing.assert_array_equal(_field(L, k, nz * nx * ny), ref)
    L.fteik_batch3d_free()


def test_station_tables_driver_matches_host_tables():
    """locate_setTravelTimeFieldsFromStations3d (the reciprocal
    station-as-source driver, fteik.c:186-235): every station's field on
    the 3D solver's grid and model equals the oracle's single-source solve,
    and the tables it hands to the locator give the same locations as
    tables uploaded from the host."""
    from tests.gpu_helpers import GPU3D
    nz, nx, ny, h = 21, 24, 22, 100.0
    vel = hetero(nz, nx, ny, h, 97)
    g = GPU3D(nz, nx, ny, h, nsweep=2, eps=5.0)
    g.set_cells(vel)
    st = np.array([[0.0, 3.3, 4.4], [0.0, 17.9, 2.1], [0.0, 9.5, 19.6]]) * h
    zs, xs, ys = (np.ascontiguousarray(st[:, i]) for i in range(3))
    ngrd = nz * nx * ny
    nev = 8
    L = lib()
    ierr = c_int(1)
    refs = [oracle3d(vel, nz, nx, ny, h, zs[k], xs[k], ys[k], 2, 5.0)[0]
            for k in range(3)]
    rng = np.random.default_rng(9)
    cells = rng.integers(0, ngrd, nev)
    picks = np.array([[r[c] + 4.0 for r in refs] for c in cells])
    o2t = np.array([2, 3, 4], dtype=np.int32)
    w = np.ones(3)
    results = []
    for source in ("stations", "host"):
        L.locate_initialize(nev, 4, ngrd, byref(ierr))
        assert ierr.value == 0
        L.locate_setReferenceCompatibility(0, byref(ierr))
        if source == "stations":
            L.locate_setTravelTimeFieldsFromStations3d(3, _d(zs), _d(xs),
                                                       _d(ys), 2, 5.0, 0, 2,
                                                       byref(ierr))
            assert ierr.value == 0
            for k in range(3):
                np.testing.assert_array_equal(_field(L, k, ngrd), refs[k])
        else:
            for k in range(3):
                L.locate_setTravelTimeField64f(ngrd, k + 2, _d(refs[k]),
                                               byref(ierr))
                assert ierr.value == 0
        one = np.ascontiguousarray(refs[0])
        L.locate_setTravelTimeField64f(ngrd, 1, _d(one), byref(ierr))
        for e in range(nev):
            p = np.ascontiguousarray(picks[e])
            L.locate_setObservation64f(e + 1, 3, False, 0.0, _i(o2t), _d(p),
                                       _d(w), byref(ierr))
        ev = np.arange(1, nev + 1, dtype=np.int32)
        out = (np.zeros(nev, dtype=np.int32), np.zeros(nev), np.zeros(nev))
        L.locate_locateEvents(nev, _i(ev), _i(out[0]), _d(out[1]),
                              _d(out[2]), byref(ierr))
        assert ierr.value == 0
        results.append(out)
        L.locate_setReferenceCompatibility(1, byref(ierr))
        L.locate_finalize()
    for a, b in zip(*results):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(results[0][0], cells + 1)
    L.fteik_batch3d_free()
