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


@pytest.mark.parametrize("tile", [7, 8, 9, 10, 11, 12, 13])
def test_mailbox_batch_order(lib, tile):
    """Mailbox tiles cover the iteration box (flipped columns 1 .. n-1)."""
    bx, by = TILES[tile]
    dims = np.array(DIMS, dtype=np.int32).ravel()
    f = lib.fteik_k_batch_mb_map
    f.restype = ctypes.c_int
    ip = ctypes.POINTER(ctypes.c_int)
    n = f(tile, len(DIMS), dims.ctypes.data_as(ip), None, 0)
    assert n > 0
    m = np.zeros(2 * n, dtype=np.int32)
    assert f(tile, len(DIMS), dims.ctypes.data_as(ip), m.ctypes.data_as(ip),
             n - 1) == -1                       # capacity is checked
    assert f(tile, len(DIMS), dims.ctypes.data_as(ip), m.ctypes.data_as(ip),
             n) == n
    entries = [tuple(e) for e in m.reshape(-1, 2)]

    def ntile(g):
        _, nx, ny = DIMS[g]
        return (nx - 2 + bx) // bx, (ny - 2 + by) // by
    _check_order(entries, DIMS, ntile)
    # the first diagonal of every grid leads
    assert [g for g, _ in entries[:len(DIMS)]] == list(range(len(DIMS)))
    # flag tiles are not mailbox tiles
    assert f(6, len(DIMS), dims.ctypes.data_as(ip), None, 0) == -1
