"""locate_locateEventsNoAlloc (the allocation-free batched grid search)
against locate_locateEvents through the C-ABI: bit-identical optimum cells,
origin times and objectives on two calls in a row, the second with fewer
events, plus the host-side size limit."""
from ctypes import byref, c_int

import numpy as np
import pytest

from tests.gpu_helpers import lib, _d, _i

pytestmark = pytest.mark.gpu


def _case(seed, nz, nx, ny, nsta, nev):
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
    picks = tables[:, true].T + rng.uniform(0, 50, nev)[:, None] + \
        rng.normal(0.0, 0.02, (nev, nsta))
    wts = rng.uniform(0.5, 1.5, (nev, nsta))
    wts[rng.random((nev, nsta)) < 0.15] = 0.0
    wts[:, 0] = 1.0
    return tables, picks, wts, ngrd


def _setup(L, tables, picks, wts, ngrd):
    nsta, nev = tables.shape[0], picks.shape[0]
    ierr = c_int(1)
    L.locate_initialize(nev, nsta, ngrd, byref(ierr))
    assert ierr.value == 0
    for s in range(nsta):
        t = np.ascontiguousarray(tables[s])
        L.locate_setTravelTimeField64f(ngrd, s + 1, _d(t), byref(ierr))
        assert ierr.value == 0
    o2t = np.arange(1, nsta + 1, dtype=np.int32)
    for e in range(nev):
        p = np.ascontiguousarray(picks[e])
        w = np.ascontiguousarray(wts[e])
        L.locate_setObservation64f(e + 1, nsta, False, 0.0, _i(o2t), _d(p),
                                   _d(w), byref(ierr))
        assert ierr.value == 0


def _run(L, fn, ev):
    n = ev.size
    opt = np.zeros(n, dtype=np.int32)
    t0 = np.zeros(n)
    ob = np.zeros(n)
    ierr = c_int(1)
    fn(n, _i(ev), _i(opt), _d(t0), _d(ob), byref(ierr))
    assert ierr.value == 0
    return opt, t0, ob


@pytest.mark.parametrize("nsta", [12, 9])
def test_noalloc_matches_allocating_entry_point(nsta):
    # nsta = 9 takes the one-pass-per-event path (no specialised kernel)
    nev = 700 if nsta == 12 else 40
    tables, picks, wts, ngrd = _case(31 + nsta, 21, 25, 19, nsta, nev)
    L = lib()
    _setup(L, tables, picks, wts, ngrd)
    try:
        rng = np.random.default_rng(5)
        first = rng.permutation(np.arange(1, nev + 1)).astype(np.int32)
        second = first[: nev // 3][::-1].copy()     # smaller, reordered
        for ev in (first, second):
            print("noalloc: nsta=%d events=%d" % (nsta, ev.size), flush=True)
            want = _run(L, L.locate_locateEvents, ev)
            got = _run(L, L.locate_locateEventsNoAlloc, ev)
            for a, b in zip(want, got):
                assert np.array_equal(a, b)
    finally:
        L.locate_finalize()


def test_noalloc_rejects_oversized_calls():
    tables, picks, wts, ngrd = _case(3, 9, 8, 7, 4, 5)
    L = lib()
    _setup(L, tables, picks, wts, ngrd)
    try:
        n = 4097
        ev = np.ones(n, dtype=np.int32)
        opt = np.zeros(n, dtype=np.int32)
        t0 = np.zeros(n)
        ob = np.zeros(n)
        ierr = c_int(0)
        L.locate_locateEventsNoAlloc(n, _i(ev), _i(opt), _d(t0), _d(ob),
                                     byref(ierr))
        assert ierr.value == 1
        # a valid call afterwards still works
        ev = np.arange(1, 6, dtype=np.int32)
        got = _run(L, L.locate_locateEventsNoAlloc, ev)
        want = _run(L, L.locate_locateEvents, ev)
        for a, b in zip(want, got):
            assert np.array_equal(a, b)
    finally:
        L.locate_finalize()
