"""Workgroup order of the batched sweeps (host bookkeeping, CPU only).

A batched sweep is one launch larger than the GPU, so it relies on
in-order dispatch: every tile must come after the tiles it waits for (its
upwind neighbours along flipped x and y in its own grid), and every
(grid, tile) pair must appear exactly once.
"""
import ctypes
import os

import numpy as np
import pytest

import fteik_amd

TILES = {  # tile id -> (bx, by), as in the library's tile table
    0: (16, 16), 3: (8, 8), 6: (16, 8),
    7: (16, 16), 8: (16, 8), 9: (8, 8), 10: (16, 16), 11: (16, 8),
    12: (8, 8), 13: (16, 8),
}
DIMS = [(23, 19, 17), (40, 37, 33), (17, 50, 45), (31, 18, 47),
        (65, 65, 65), (5, 3, 129)]


@pytest.fixture(scope="module")
def lib():
    if not os.path.isfile(fteik_amd.library_path()):
        pytest.fail("libfteik_shared.so is not built")
    return ctypes.CDLL(fteik_amd.library_path())


def _check_order(entries, dims, ntile_of):
    """entries: list of (grid, tile) in workgroup order; ntile_of(g) ->
    (nbx, nby).  Each pair once; upwind neighbours earlier."""
    pos = {}
    for w, (g, t) in enumerate(entries):
        assert (g, t) not in pos
        pos[(g, t)] = w
    for g in range(len(dims)):
        nbx, nby = ntile_of(g)
        for bx in range(nbx):
            for by in range(nby):
                w = pos.pop((g, bx * nby + by))
                if bx > 0:
                    assert pos.get((g, (bx - 1) * nby + by), -1) < w
                if by > 0:
                    assert pos.get((g, bx * nby + by - 1), -1) < w
    assert not pos


SRCS = [(8, 11, 6), (31, 4, 21), (1, 48, 2), (13, 16, 45), (33, 33, 33),
        (2, 1, 127)]                       # 1-based cell indices


def _box(sweep, n, src):
    """Initialization box of one axis, flipped 0-based (inclusive), as the
    reference's init masks give it (solver3D.F90:493-565)."""
    tz = [1, 1, 1, 1, -1, -1, -1, -1]
    tx = [1, -1, 1, -1, 1, -1, 1, -1]
    ty = [1, 1, -1, -1, 1, 1, -1, -1]
    out = []
    for sg, nn, si in zip((tz[sweep - 1], tx[sweep - 1], ty[sweep - 1]), n,
                          src):
        a0, a1 = (max(2, si), nn) if sg > 0 else (1, si + 1)
        lo, hi = (a0 - 1, a1 - 1) if sg > 0 else (nn - a1, nn - a0)
        out.append((lo, hi))
    return out


def _map(lib, tile, sweep):
    dims = np.array(DIMS, dtype=np.int32).ravel()
    srcs = np.array(SRCS, dtype=np.int32).ravel()
    f = lib.fteik_k_batch_mb_map
    f.restype = ctypes.c_int
    ip = ctypes.POINTER(ctypes.c_int)
    args = (tile, len(DIMS), dims.ctypes.data_as(ip), srcs.ctypes.data_as(ip),
            sweep)
    n = f(*args, None, 0)
    assert n >= 0
    m = np.zeros(2 * max(n, 1), dtype=np.int32)
    if n > 0:
        assert f(*args, m.ctypes.data_as(ip), n - 1) == -1   # capacity
    assert f(*args, m.ctypes.data_as(ip), n) == n
    return [tuple(e) for e in m[:2 * n].reshape(-1, 2)]


@pytest.mark.parametrize("tile", [7, 8, 9, 10, 11, 12, 13])
def test_mailbox_batch_order(lib, tile):
    """Iteration sweeps: tiles cover flipped columns 1 .. n-1."""
    bx, by = TILES[tile]
    entries = _map(lib, tile, 0)

    def ntile(g):
        _, nx, ny = DIMS[g]
        return (nx - 2 + bx) // bx, (ny - 2 + by) // by
    _check_order(entries, DIMS, ntile)
    # the first diagonal of every grid leads
    assert [g for g, _ in entries[:len(DIMS)]] == list(range(len(DIMS)))
    # flag tiles are not mailbox tiles
    dims = np.array(DIMS, dtype=np.int32).ravel()
    ip = ctypes.POINTER(ctypes.c_int)
    assert lib.fteik_k_batch_mb_map(6, len(DIMS), dims.ctypes.data_as(ip),
                                    None, 0, None, 0) == -1


@pytest.mark.parametrize("tile", [7, 9, 13])
@pytest.mark.parametrize("sweep", range(1, 9))
def test_mailbox_batch_init_order(lib, tile, sweep):
    """Initialization sweeps: each grid's tiles cover its source box (empty
    boxes give no entries)."""
    bx, by = TILES[tile]
    entries = _map(lib, tile, sweep)

    def ntile(g):
        (z0, z1), (x0, x1), (y0, y1) = _box(sweep, DIMS[g], SRCS[g])
        if z0 > z1 or x0 > x1 or y0 > y1:
            return 0, 0
        return (x1 - x0 + bx) // bx, (y1 - y0 + by) // by
    _check_order(entries, DIMS, ntile)
