"""GPU parity of the grid-search locator (locate.F90) through the C-ABI.

Reference compatibility mode (default) reproduces the reference's results
on its own fixtures exactly, carried-over origin-time state and zero-padded
minimum search included; fresh mode and the batched entry point equal the
oracle's fresh-state search bit for bit.
"""
from ctypes import byref, c_double, c_int

import numpy as np
import pytest

from oracle import oracle as orc
from tests.gpu_helpers import lib, _d, _i

pytestmark = pytest.mark.gpu


def _setup(nev, ntf, ngrd, tables, picks, wts, obs2tf=None):
    L = lib()
    ierr = c_int(1)
    L.locate_initialize(nev, ntf, ngrd, byref(ierr))
    assert ierr.value == 0
    for s in range(ntf):
        t = np.ascontiguousarray(tables[s], dtype=np.float64)
        L.locate_setTravelTimeField64f(ngrd, s + 1, _d(t), byref(ierr))
        assert ierr.value == 0
    for e in range(nev):
        o2t = np.arange(1, ntf + 1, dtype=np.int32) if obs2tf is None \
            else np.ascontiguousarray(obs2tf[e], dtype=np.int32)
        p = np.ascontiguousarray(picks[e], dtype=np.float64)
        w = np.ascontiguousarray(wts[e], dtype=np.float64)
        L.locate_setObservation64f(e + 1, p.size, False, 0.0, _i(o2t), _d(p),
                                   _d(w), byref(ierr))
        assert ierr.value == 0
    return L


def _locate(L, ev):
    opt, t0, obj = c_int(), c_double(), c_double()
    L.locate_locateEvent(ev, byref(opt), byref(t0), byref(obj))
    return opt.value, t0.value, obj.value


@pytest.mark.parametrize("tag", ["small", "mid"])
def test_reference_mode_matches_reference_fixtures(fixtures, tag):
    nz, nx, ny, h, nsta, nev = (int(v) for v in fixtures["loc_%s_cfg" % tag])
    ngrd = nz * nx * ny
    L = _setup(nev, nsta, ngrd, fixtures["loc_%s_tables" % tag],
               fixtures["loc_%s_picks" % tag], fixtures["loc_%s_wts" % tag])
    res = fixtures["loc_%s_result" % tag]
    for e in range(nev):
        opt, t0, ob = _locate(L, e + 1)
        assert opt == int(res[e, 0]), (e, opt, res[e])
        assert t0 == res[e, 1] and ob == res[e, 2]
    L.locate_finalize()


def _random_case(seed, nz, nx, ny, nsta, nev, drop=0.0):
    rng = np.random.default_rng(seed)
    h = 100.0
    ngrd = nz * nx * ny
    iz, ix, iy = np.meshgrid(np.arange(nz), np.arange(nx), np.arange(ny),
                             indexing="ij")
    zz = (iz * h).transpose(2, 1, 0).ravel()
    xx = (ix * h).transpose(2, 1, 0).ravel()
    yy = (iy * h).transpose(2, 1, 0).ravel()
    sx = rng.uniform(0, (nx - 1) * h, nsta)
    sy = rng.uniform(0, (ny - 1) * h, nsta)
    tables = np.sqrt((xx[None] - sx[:, None]) ** 2 +
                     (yy[None] - sy[:, None]) ** 2 + zz[None] ** 2) / 3500.0
    true = rng.integers(0, ngrd, nev)
    t0 = rng.uniform(0.0, 100.0, nev)
    picks = tables[:, true].T + t0[:, None] + \
        rng.normal(0.0, 0.01, (nev, nsta))
    wts = rng.uniform(0.5, 1.5, (nev, nsta))
    if drop:
        wts[rng.random((nev, nsta)) < drop] = 0.0
        wts[:, 0] = 1.0
    return tables, picks, wts, ngrd


@pytest.mark.parametrize("nsta", [5, 12, 48])
def test_fresh_and_batched_match_oracle(nsta):
    nz, nx, ny, nev = 19, 23, 17, 40
    tables, picks, wts, ngrd = _random_case(nsta, nz, nx, ny, nsta, nev,
                                            drop=0.2)
    L = _setup(nev, nsta, ngrd, tables, picks, wts)
    ierr = c_int(1)
    L.locate_setReferenceCompatibility(0, byref(ierr))
    assert ierr.value == 0
    loc = orc.Locator(nev, nsta, ngrd)
    for s in range(nsta):
        loc.set_table(s + 1, tables[s])
    want = []
    for e in range(nev):
        loc.set_observation(e + 1, np.arange(1, nsta + 1), picks[e], wts[e])
        want.append(loc.locate(e + 1, accumulate_t0=False, pad_argmin=False))
    for e in range(nev):
        assert _locate(L, e + 1) == want[e], e
    ev = np.arange(1, nev + 1, dtype=np.int32)[::-1].copy()
    opt = np.zeros(nev, dtype=np.int32)
    t0 = np.zeros(nev)
    ob = np.zeros(nev)
    L.locate_locateEvents(nev, _i(ev), _i(opt), _d(t0), _d(ob), byref(ierr))
    assert ierr.value == 0
    for q, e in enumerate(ev):
        assert (opt[q], t0[q], ob[q]) == want[e - 1], (q, e)
    L.locate_setReferenceCompatibility(1, byref(ierr))
    L.locate_finalize()


@pytest.mark.parametrize("nev", [1, 2, 8, 9])
def test_few_events_match_oracle(nev):
    """Batched searches of at most 8 events post one key per workgroup
    (a kernel variant of their own); 9 events take the many-event kernel.
    Both equal the oracle bit for bit (locate.F90:620-744, MINLOC :122)."""
    nsta = 12
    tables, picks, wts, ngrd = _random_case(21, 41, 37, 29, nsta, nev)
    L = _setup(nev, nsta, ngrd, tables, picks, wts)
    ierr = c_int(1)
    L.locate_setReferenceCompatibility(0, byref(ierr))
    loc = orc.Locator(nev, nsta, ngrd)
    for s in range(nsta):
        loc.set_table(s + 1, tables[s])
    ev = np.arange(1, nev + 1, dtype=np.int32)
    opt = np.zeros(nev, dtype=np.int32)
    t0 = np.zeros(nev)
    ob = np.zeros(nev)
    L.locate_locateEvents(nev, _i(ev), _i(opt), _d(t0), _d(ob), byref(ierr))
    assert ierr.value == 0
    for e in range(nev):
        loc.set_observation(e + 1, np.arange(1, nsta + 1), picks[e], wts[e])
        assert (opt[e], t0[e], ob[e]) == loc.locate(e + 1, False, False)
    L.locate_setReferenceCompatibility(1, byref(ierr))
    L.locate_finalize()


def test_table_count_without_specialised_kernel():
    """ntf = 9 has no register-resident batch kernel: one pass per event."""
    nsta, nev = 9, 6
    tables, picks, wts, ngrd = _random_case(9, 11, 13, 10, nsta, nev)
    L = _setup(nev, nsta, ngrd, tables, picks, wts)
    ierr = c_int(1)
    L.locate_setReferenceCompatibility(0, byref(ierr))
    loc = orc.Locator(nev, nsta, ngrd)
    for s in range(nsta):
        loc.set_table(s + 1, tables[s])
    ev = np.arange(1, nev + 1, dtype=np.int32)
    opt = np.zeros(nev, dtype=np.int32)
    t0 = np.zeros(nev)
    ob = np.zeros(nev)
    L.locate_locateEvents(nev, _i(ev), _i(opt), _d(t0), _d(ob), byref(ierr))
    assert ierr.value == 0
    for e in range(nev):
        loc.set_observation(e + 1, np.arange(1, nsta + 1), picks[e], wts[e])
        assert (opt[e], t0[e], ob[e]) == loc.locate(e + 1, False, False)
    L.locate_setReferenceCompatibility(1, byref(ierr))
    L.locate_finalize()


def test_float_tables_and_error_paths():
    nsta, nev = 4, 3
    tables, picks, wts, ngrd = _random_case(4, 8, 9, 10, nsta, nev)
    L = lib()
    ierr = c_int(1)
    L.locate_initialize(nev, nsta, ngrd, byref(ierr))
    assert ierr.value == 0
    t32 = np.ascontiguousarray(tables[0], dtype=np.float32)
    import ctypes
    L.locate_setTravelTimeField32f(ngrd, 1, t32.ctypes.data_as(
        ctypes.POINTER(ctypes.c_float)), byref(ierr))
    assert ierr.value == 0
    bad = np.ascontiguousarray(tables[1])
    L.locate_setTravelTimeField64f(ngrd, nsta + 1, _d(bad), byref(ierr))
    assert ierr.value == 1
    L.locate_setTravelTimeField64f(ngrd - 1, 2, _d(bad), byref(ierr))
    assert ierr.value == 1
    o2t = np.array([0, 1, 2, 3], dtype=np.int32)       # 0 is invalid
    p = np.ascontiguousarray(picks[0])
    w = np.ascontiguousarray(wts[0])
    L.locate_setObservation64f(1, 4, False, 0.0, _i(o2t), _d(p), _d(w),
                               byref(ierr))
    assert ierr.value == 1
    o2t = np.array([1, 2, 3, 4], dtype=np.int32)
    L.locate_setObservation64f(nev + 1, 4, False, 0.0, _i(o2t), _d(p), _d(w),
                               byref(ierr))
    assert ierr.value == 1
    L.locate_finalize()


def test_statics():
    """locate.F90:748-782: per observation slot, the float sum over events
    of w * (tobs - (tt + t0)), accumulated in event order."""
    ldo, nobs, nev = 6, 5, 4
    rng = np.random.default_rng(1)
    tt = rng.uniform(1, 5, (nev, ldo)).astype(np.float32)
    t0 = rng.uniform(0, 2, nev).astype(np.float32)
    stat = rng.uniform(-0.1, 0.1, ldo).astype(np.float32)
    tobs = (tt + t0[:, None] + stat[None, :]).astype(np.float32)
    wts = rng.uniform(0.5, 1.0, (nev, ldo)).astype(np.float32)
    out = np.zeros(ldo, dtype=np.float32)
    import ctypes
    fp = ctypes.POINTER(ctypes.c_float)
    L = lib()
    a, b, c = c_int(ldo), c_int(nobs), c_int(nev)
    L.locate_computeStatics32fF(byref(a), byref(b), byref(c),
                                tobs.ctypes.data_as(fp), tt.ctypes.data_as(fp),
                                t0.ctypes.data_as(fp), wts.ctypes.data_as(fp),
                                out.ctypes.data_as(fp))
    want = np.zeros(nobs, dtype=np.float32)
    for e in range(nev):
        test = tt[e, :nobs] + t0[e]
        want = want + wts[e, :nobs] * (tobs[e, :nobs] - test)
    np.testing.assert_array_equal(out[:nobs], want)
    # the residuals are the planted statics times the summed weights
    np.testing.assert_allclose(out[:nobs], stat[:nobs] *
                               wts[:, :nobs].sum(axis=0), atol=1e-5)
