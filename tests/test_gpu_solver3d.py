"""GPU parity of the 3D level-set solver against the oracle.

Nodes of one level never read each other, so any schedule that respects
the levels yields the reference's field bit for bit: the FP64 device path
must equal the oracle exactly (the specification allows 1e-12 relative;
we ask for equality).  All calls go through the C-ABI.
"""
import ctypes
from ctypes import byref, c_int

import numpy as np
import pytest

from oracle import oracle as orc
from tests.gpu_helpers import GPU3D, grid9, hetero, lib, oracle3d, _d, _i

pytestmark = pytest.mark.gpu


def _announce(*a):
    print("case:", *a, flush=True)


# ------------------------------------------------------------ stencils
def test_local_solvers_match_reference_fixtures(fixtures):
    """Device noinit/init stencils on the reference's own 1000 node inputs
    (localSolver.F90:384-751, fixtures from the reference build)."""
    nz, nx, ny, h, zs, xs, ys, eps = fixtures["ls3d_grid"]
    nz, nx, ny = int(nz), int(nx), int(ny)
    g = GPU3D(nz, nx, ny, h, nsweep=2, eps=eps)
    g.set_cells(fixtures["ls3d_vel"])
    g.set_sources(zs, xs, ys)
    L = lib()
    ierr = c_int(1)
    L.fteik_amd_prepareSource3d(1, byref(ierr))
    assert ierr.value == 0
    rows = np.ascontiguousarray(fixtures["ls3d_in"])
    nodes = np.ascontiguousarray(fixtures["ls3d_node"], dtype=np.int32)
    n = rows.shape[0]
    out_n, out_i = np.zeros(n), np.zeros(n)
    L.fteik_amd_evalLocalSolver3d(n, _d(rows), _i(nodes), _d(out_n),
                                  _d(out_i), byref(ierr))
    assert ierr.value == 0
    np.testing.assert_array_equal(out_n, fixtures["ls3d_noinit"])
    np.testing.assert_array_equal(out_i, fixtures["ls3d_init"])


# ------------------------------------------------------------- solves
CASES = [
    # nz, nx, ny, h, (zs, xs, ys) in cells, nsweep, eps, seed
    (23, 19, 17, 15.0, (7.7, 10.2, 5.5), 2, 4.0, 3),
    (40, 37, 33, 12.0, (30.25, 3.5, 20.0), 2, 5.0, 4),     # >1 block per axis
    (17, 50, 45, 10.0, (0.0, 47.2, 1.1), 1, 3.0, 5),       # z at the surface
    (33, 33, 33, 10.0, (16.0, 16.0, 16.0), 0, 5.0, 6),     # init sweeps only
    (31, 18, 47, 20.0, (12.6, 15.9, 44.3), 3, 2.0, 7),     # near far corner
]


@pytest.mark.parametrize("schedule", [0, 1])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_solve_matches_oracle(case, schedule):
    nz, nx, ny, h, (zc, xc, yc), nsweep, eps, seed = CASES[case]
    _announce("solve3d", CASES[case], "schedule", schedule)
    vel = hetero(nz, nx, ny, h, seed)
    zs, xs, ys = zc * h, xc * h, yc * h
    ref, _, _ = oracle3d(vel, nz, nx, ny, h, zs, xs, ys, nsweep, eps)
    g = GPU3D(nz, nx, ny, h, nsweep=nsweep, eps=eps, schedule=schedule)
    g.set_cells(vel)
    g.set_sources(zs, xs, ys)
    assert g.solve(1) == 0
    np.testing.assert_array_equal(g.field(), ref)
    assert ref.max() < orc.HUGE


def test_linear_gradient_101():
    """The 101^3 workload: h = 10 m, v(z) = 3000 + 0.5 z, source at the
    centre, 20 seeded surface receivers, 2 iterations.  The field equals
    the oracle's bit for bit, receivers too, and stays within first-order
    accuracy of the closed form (analytic.F90:456-528)."""
    n, h = 101, 10.0
    _announce("linear gradient", n)
    zc = (np.arange(n - 1) + 0.5) * h
    vel = np.tile(3000.0 + 0.5 * zc, (n - 1) * (n - 1))
    zs = xs = ys = 50 * h
    ref, slow, src = oracle3d(vel, n, n, n, h, zs, xs, ys, 2, 5.0)
    rng = np.random.default_rng(101)
    xr = rng.uniform(0.0, (n - 1) * h, 20)
    yr = rng.uniform(0.0, (n - 1) * h, 20)
    zr = np.zeros(20)
    g = GPU3D(n, n, n, h, nsweep=2, eps=5.0)
    g.set_cells(vel)
    g.set_sources(zs, xs, ys)
    g.set_receivers(zr, xr, yr)
    assert g.solve(1) == 0
    tt = g.field()
    np.testing.assert_array_equal(tt, ref)
    ttr_ref = orc.receivers(ref, grid9(n, n, n, h), zr, xr, yr)
    np.testing.assert_array_equal(g.receivers(), ttr_ref)
    vbot = 3000.0 + 0.5 * (n - 1) * h
    a = orc.xfast_to_natural(
        orc.analytic_gradient(n, n, n, h, h, h, 0.0, 0.0, 0.0, zs, xs, ys,
                              3000.0, vbot), n, n, n)
    rel = np.abs(tt - a) / np.maximum(a, 1e-12)
    assert np.median(rel) < 2e-3 and rel.max() < 1e-2


def test_getters_orders_and_cells():
    nz, nx, ny, h = 21, 18, 14, 25.0
    vel = hetero(nz, nx, ny, h, 11)
    zs, xs, ys = 4.4 * h, 9.1 * h, 6.6 * h
    ref, _, _ = oracle3d(vel, nz, nx, ny, h, zs, xs, ys, 2, 5.0)
    g = GPU3D(nz, nx, ny, h)
    g.set_cells(vel)
    g.set_sources(zs, xs, ys)
    assert g.solve(1) == 0
    for order in (1, 2):
        np.testing.assert_array_equal(
            g.field(order), orc.field_reorder3d(ref, nz, nx, ny, order))
    for order in (0, 1, 2):
        np.testing.assert_array_equal(
            g.cells(order), orc.field_cell3d(ref, nz, nx, ny, order))


@pytest.mark.parametrize("order", [0, 1, 2])
def test_nodal_model_orders(order):
    """Nodal velocities (model.F90:311-440, quirks kept) in each order."""
    nz, nx, ny, h = 15, 13, 12, 20.0
    rng = np.random.default_rng(order)
    vn = 3000.0 + 1000.0 * rng.random(nz * nx * ny)
    slow = orc.nodal_slowness(vn, nz, nx, ny, order)
    zs, xs, ys = 5.5 * h, 6.2 * h, 3.3 * h
    src = orc.source_locate(grid9(nz, nx, ny, h), zs, xs, ys, slow)
    ref = orc.solve3d(slow, nz, nx, ny, h, h, h, src, 2, 5.0, schedule=1)
    g = GPU3D(nz, nx, ny, h)
    ierr = c_int(1)
    v = np.ascontiguousarray(vn)
    g.L.fteik_solver3d_setNodalVelocityModel64f(v.size, order, _d(v),
                                                byref(ierr))
    assert ierr.value == 0
    g.set_sources(zs, xs, ys)
    assert g.solve(1) == 0
    np.testing.assert_array_equal(g.field(), ref)


@pytest.mark.parametrize("order", [1, 2])
def test_cell_model_orders(order):
    """Cell velocities in XYZ (order 1) and ZYX (order 2) order
    (model.F90:469-559), the path python/fteik3d.py takes with ZYX."""
    nz, nx, ny, h = 19, 16, 14, 25.0
    vel_nat = hetero(nz, nx, ny, h, 50 + order)
    v = vel_nat.reshape(ny - 1, nx - 1, nz - 1)
    vo = (v.transpose(2, 0, 1) if order == 1 else v.transpose(1, 0, 2))
    vo = np.ascontiguousarray(vo).ravel()
    slow = orc.cell_slowness(vo, nz, nx, ny, order)
    np.testing.assert_array_equal(slow, 1.0 / vel_nat)
    zs, xs, ys = 6.6 * h, 3.3 * h, 9.9 * h
    src = orc.source_locate(grid9(nz, nx, ny, h), zs, xs, ys, slow)
    ref = orc.solve3d(slow, nz, nx, ny, h, h, h, src, 2, 5.0, schedule=1)
    g = GPU3D(nz, nx, ny, h)
    ierr = c_int(1)
    g.L.fteik_solver3d_setCellVelocityModel64f(vo.size, order, _d(vo),
                                               byref(ierr))
    assert ierr.value == 0
    g.set_sources(zs, xs, ys)
    assert g.solve(1) == 0
    np.testing.assert_array_equal(g.field(), ref)


def test_index_helpers_round_trip():
    """testing/tests.c:262-302: grid2indexF and index2gridF invert each
    other on the 34 x 10 x 21 grid the solver was initialized with; the
    0-based shim (fwrapper.c:8-14) and the cell index (model.F90:827-836)
    by their closed forms; index2gridF follows the geometry of the solver
    initialized last (ny = 1 for the 2D solver) and, with none, reports
    ierr = 3."""
    L = lib()
    nz, nx, ny = 34, 10, 21
    g = GPU3D(nz, nx, ny, 10.0)
    i1, j1, k1, ierr = c_int(), c_int(), c_int(), c_int()
    nzx = nz * nx
    for k in range(1, ny + 1):
        for j in range(1, nx + 1):
            for i in range(1, nz + 1):
                idx = L.fteik_model_grid2indexF(i, j, k, nz, nzx)
                assert idx == (k - 1) * nzx + (j - 1) * nz + i
                assert L.fteik_model_grid2index(i - 1, j - 1, k - 1, nz,
                                                nzx) == idx - 1
                L.fteik_model_index2gridF(idx, byref(i1), byref(j1),
                                          byref(k1), byref(ierr))
                assert ierr.value == 0
                assert (i1.value, j1.value, k1.value) == (i, j, k)
    assert L.fteik_model_velGrid2indexF(3, 4, 5, nz - 1, (nz - 1) * (nx - 1)) \
        == 4 * (nz - 1) * (nx - 1) + 3 * (nz - 1) + 3
    # past the last node: k out of range
    L.fteik_model_index2gridF(nzx * ny + 1, byref(i1), byref(j1), byref(k1),
                              byref(ierr))
    assert ierr.value == 1
    # the 2D solver's geometry: ny = 1
    L.fteik_solver2d_initialize(12, 9, 0.0, 0.0, 5.0, 5.0, 1, 3.0, 0.0, 0,
                                byref(ierr))
    assert ierr.value == 0
    L.fteik_model_index2gridF(12 * 9, byref(i1), byref(j1), byref(k1),
                              byref(ierr))
    assert ierr.value == 0 and (i1.value, j1.value, k1.value) == (12, 9, 1)
    L.fteik_model_index2gridF(12 * 9 + 1, byref(i1), byref(j1), byref(k1),
                              byref(ierr))
    assert ierr.value == 1
    L.fteik_solver2d_free()
    L.fteik_model_index2gridF(5, byref(i1), byref(j1), byref(k1),
                              byref(ierr))
    assert ierr.value == 3
    del g


def test_multiple_sources_and_origin():
    """Several sources, solved one after another on the same state, with a
    non-zero grid origin (source.F90:136-190)."""
    nz, nx, ny, h = 26, 22, 19, 30.0
    origin = (100.0, -250.0, 40.0)
    vel = hetero(nz, nx, ny, h, 21)
    srcs = [(3.3, 5.0, 7.7), (20.2, 1.0, 0.4), (12.0, 12.0, 9.0)]
    g = GPU3D(nz, nx, ny, h, origin=origin)
    g.set_cells(vel)
    z = [origin[0] + s[0] * h for s in srcs]
    x = [origin[1] + s[1] * h for s in srcs]
    y = [origin[2] + s[2] * h for s in srcs]
    g.set_sources(z, x, y)
    for k in range(len(srcs)):
        ref, _, _ = oracle3d(vel, nz, nx, ny, h, z[k], x[k], y[k], 2, 5.0,
                             origin=origin)
        assert g.solve(k + 1) == 0
        np.testing.assert_array_equal(g.field(), ref)


def test_mixed_precision_converges_to_fp64():
    """FP32 iterations, promotion with an upward margin, then FP64
    iterations until nothing changes: the result is the converged FP64
    field to <= 1e-6 relative (the tolerance the mixed mode is specified
    to); the FP64 reference runs 16 iterations, far past convergence."""
    nz, nx, ny, h = 48, 45, 41, 20.0
    vel = hetero(nz, nx, ny, h, 31)
    zs, xs, ys = 20.3 * h, 11.1 * h, 30.7 * h
    g = GPU3D(nz, nx, ny, h, nsweep=16)
    g.set_cells(vel)
    g.set_sources(zs, xs, ys)
    assert g.solve(1) == 0
    t64 = g.field()
    g8 = GPU3D(nz, nx, ny, h, nsweep=2, precision=1, conv=0.0)
    g8.set_cells(vel)
    g8.set_sources(zs, xs, ys)
    assert g8.solve(1) == 0
    tm = g8.field()
    _, corr, _ = g8.info()
    assert 1 <= corr < 16          # converged before the cap
    rel = np.abs(tm - t64) / np.maximum(t64, 1e-12)
    print("mixed: %d FP64 iterations, max rel err %.3g" % (corr, rel.max()))
    assert rel.max() <= 1e-6, rel.max()


def test_error_paths():
    L = lib()
    ierr = c_int(0)
    g = GPU3D(12, 11, 10, 10.0)
    # no model yet
    g.set_sources(50.0, 50.0, 50.0)
    L.fteik_solver3d_solveSourceLSM(1, byref(ierr))
    assert ierr.value == 1
    g.set_cells(np.full(11 * 10 * 9, 3000.0))
    # source number out of range
    L.fteik_solver3d_solveSourceLSM(2, byref(ierr))
    assert ierr.value == 1
    # source in the last cell along z (the reference reads out of bounds)
    g.set_sources(10.95 * 10.0, 50.0, 50.0)
    L.fteik_solver3d_solveSourceLSM(1, byref(ierr))
    assert ierr.value == 1
    # field requested before any successful solve
    tt = np.zeros(12 * 11 * 10)
    L.fteik_solver3d_getTravelTimeField64f(tt.size, 0, _d(tt), byref(ierr))
    assert ierr.value == 1
    # negative velocity
    bad = np.full(11 * 10 * 9, 3000.0)
    bad[17] = -1.0
    L.fteik_solver3d_setVelocityModel64f(bad.size, _d(bad), byref(ierr))
    assert ierr.value == 1
    # velocities above the supported range, and NaN
    for v in (2e30, np.nan):
        bad = np.full(11 * 10 * 9, 3000.0)
        bad[5] = v
        L.fteik_solver3d_setVelocityModel64f(bad.size, _d(bad), byref(ierr))
        assert ierr.value == 1
    ok = np.full(11 * 10 * 9, 1e30)
    L.fteik_solver3d_setVelocityModel64f(ok.size, _d(ok), byref(ierr))
    assert ierr.value == 0
    # bad schedule / precision values
    L.fteik_solver3d_setSchedule(7, byref(ierr))
    assert ierr.value == 1
    L.fteik_solver3d_setPrecision(3, byref(ierr))
    assert ierr.value == 1
    # source outside the grid
    z = np.array([-5.0])
    L.fteik_solver3d_setSources64f(1, _d(z), _d(z), _d(z), byref(ierr))
    assert ierr.value == 1


def test_stencil_sqrt_is_correctly_rounded():
    """The iteration stencil's square root (used by its 2D operators)
    equals IEEE sqrt on its stated range: log-uniform arguments over
    [2^-767, 2^1023], exact squares, their floating-point neighbours (where
    rounding is hardest) and the range ends."""
    L = lib()
    rng = np.random.default_rng(17)
    x = [np.exp2(rng.uniform(-767.0, 1023.0, 200000))]
    r = rng.uniform(1e-100, 1e100, 20000)
    sq = r * r
    x += [sq, np.nextafter(sq, 0.0), np.nextafter(sq, np.inf),
          rng.integers(1, 2**26, 20000).astype(np.float64) ** 2]
    x.append(np.array([2.0 ** -767, np.finfo(np.float64).max, 1.0, 4.0,
                       2.0, 0.5]))
    x = np.ascontiguousarray(np.concatenate(x))
    x = x[np.isfinite(x) & (x >= 2.0 ** -767)]
    y = np.zeros_like(x)
    ierr = c_int(1)
    L.fteik_amd_sqrtStencil64f(x.size, _d(x), _d(y), byref(ierr))
    assert ierr.value == 0
    np.testing.assert_array_equal(y, np.sqrt(x))


def test_grid_spacing_limit():
    L = lib()
    ierr = c_int(0)
    L.fteik_solver3d_initialize(9, 8, 7, 0.0, 0.0, 0.0, 10.0, 2e30, 10.0, 2,
                                5.0, 0.0, 0, byref(ierr))
    assert ierr.value == 1
    L.fteik_solver3d_initialize(9, 8, 7, 0.0, 0.0, 0.0, 10.0, 10.0, 10.0, 2,
                                5.0, 0.0, 0, byref(ierr))
    assert ierr.value == 0


def test_graph_levels_match_oracle():
    """graph3d.F90: level pointers and (i, j, k, node) lists per sweep."""
    L = lib()
    nz, nx, ny = 9, 7, 6
    ierr = c_int(1)
    L.fteik_graph3d_initialize(nz, nx, ny, 0, byref(ierr))
    assert ierr.value == 0
    L.fteik_graph3d_makeLevelStructures(byref(ierr))
    assert ierr.value == 0
    nlev = L.fteik_graph3d_getNumberOfLevels(byref(ierr))
    assert nlev == nz + nx + ny - 2
    ngrd = nz * nx * ny
    for sweep in range(1, 9):
        ptr_ref, ijkv_ref = orc.levels3d(sweep, nz, nx, ny)
        ptr = np.zeros(nlev + 1, dtype=np.int32)
        L.fteik_graph3d_getLevelPointerF(nlev + 1, _i(ptr), byref(ierr))
        assert ierr.value == 0
        np.testing.assert_array_equal(ptr, ptr_ref)
        ijkv = np.zeros(4 * ngrd, dtype=np.int32)
        L.fteik_graph3d_getIJKVF(4 * ngrd, sweep, _i(ijkv), byref(ierr))
        assert ierr.value == 0
        np.testing.assert_array_equal(ijkv, ijkv_ref)
    mx = L.fteik_graph3d_getMaxLevelSize(byref(ierr))
    assert mx == int(np.max(np.diff(ptr_ref)))
    L.fteik_graph3d_free()


@pytest.mark.parametrize("precision", [0, 1])
def test_every_tile_shape_matches_oracle(precision):
    """Every workgroup shape and halo-exchange mode of the pipelined sweep
    (flags, mailbox, mailbox with store wave) gives the oracle's field bit
    for bit, in FP64 and in the mixed mode (which converges onto it)."""
    nz, nx, ny, h = 40, 37, 33, 12.0
    vel = hetero(nz, nx, ny, h, 4)
    zs, xs, ys = 30.25 * h, 3.5 * h, 20.0 * h
    ref, _, _ = oracle3d(vel, nz, nx, ny, h, zs, xs, ys, 4, 5.0)
    g = GPU3D(nz, nx, ny, h, nsweep=4, precision=precision)
    g.set_cells(vel)
    g.set_sources(zs, xs, ys)
    L = lib()
    t, t32, nt, ierr = c_int(), c_int(), c_int(), c_int()
    L.fteik_solver3d_setSweepOverlap(0, byref(ierr))   # one launch per sweep
    L.fteik_solver3d_getTile(byref(t), byref(t32), byref(nt), byref(ierr))
    assert ierr.value == 0 and nt.value >= 14
    tested = 0
    for tile in range(nt.value):
        L.fteik_solver3d_setTile(tile, tile, byref(ierr))
        if ierr.value != 0:
            continue
        _announce("tile", tile, "precision", precision)
        assert g.solve(1) == 0
        np.testing.assert_array_equal(g.field(), ref)
        tested += 1
    assert tested == nt.value          # every shape fits this grid


@pytest.mark.parametrize("skipping", [0, 1])
def test_tiles_launched_without_coresidency(skipping):
    """On a 129 x 257 x 257 grid the 1024 workgroups of the 8 x 8
    store-wave tiles (3 waves each) are not all resident at once: the
    sweeps launch them without cooperation, in tile order.  The field equals
    the oracle's and the cooperative 8 x 8 tiles' bit for bit, with and
    without tile skipping, also with the staged loads (tile 18, the default
    of such grids)."""
    nz, nx, ny, h = 129, 257, 257, 20.0
    vel = hetero(nz, nx, ny, h, 12)
    zs, xs, ys = 60.3 * h, 100.7 * h, 200.2 * h
    ref, _, _ = oracle3d(vel, nz, nx, ny, h, zs, xs, ys, 1, 5.0)
    g = GPU3D(nz, nx, ny, h, nsweep=1)
    g.set_cells(vel)
    g.set_sources(zs, xs, ys)
    L = lib()
    ierr = c_int(1)
    L.fteik_solver3d_setSweepOverlap(0, byref(ierr))
    L.fteik_solver3d_setTileSkipping(skipping, byref(ierr))
    t, t32, nt = c_int(), c_int(), c_int()
    L.fteik_solver3d_getTile(byref(t), byref(t32), byref(nt), byref(ierr))
    assert t.value == 18 and t32.value == 18
    # the FP64 launches below are the non-cooperative ones: a tile plane of
    # this grid does not fit the GPU at once (library's own residency check)
    for tile in (18, 12):
        assert L.fteik_k_pencil_fits(nx, ny, 0, tile) == 0
    for tile in (18, 12, 9):
        _announce("large grid tile", tile, "skipping", skipping)
        L.fteik_solver3d_setTile(tile, tile, byref(ierr))
        assert ierr.value == 0
        assert g.solve(1) == 0
        np.testing.assert_array_equal(g.field(), ref)


def test_tile_skipping_is_exact():
    """Many iterations (so later ones skip most tiles): the field equals the
    oracle's, and equals the solver's own dense result, bit for bit."""
    nz, nx, ny, h = 45, 52, 38, 15.0
    vel = hetero(nz, nx, ny, h, 8)
    zs, xs, ys = 11.2 * h, 40.4 * h, 20.7 * h
    ref, _, _ = oracle3d(vel, nz, nx, ny, h, zs, xs, ys, 8, 5.0)
    L = lib()
    ierr = c_int()
    fields = []
    for on in (1, 0):
        g = GPU3D(nz, nx, ny, h, nsweep=8)
        L.fteik_solver3d_setSweepOverlap(0, byref(ierr))
        L.fteik_solver3d_setTileSkipping(on, byref(ierr))
        g.set_cells(vel)
        g.set_sources(zs, xs, ys)
        assert g.solve(1) == 0
        fields.append(g.field())
    np.testing.assert_array_equal(fields[0], ref)
    np.testing.assert_array_equal(fields[1], ref)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("case", [
    # nz, nx, ny, h, source (z, x, y) in cells, nsweep
    (40, 37, 33, 12.0, (30.25, 3.5, 20.0), 3),
    (33, 64, 49, 10.0, (1.5, 60.2, 2.7), 2),
    (71, 20, 45, 20.0, (35.0, 10.0, 22.0), 5),
])
def test_sweep_overlap_matches_oracle(case, precision):
    """Iteration sweeps chained in one launch, each overlapping the one
    before it, for every store-wave tile shape: the field is the oracle's
    bit for bit (FP64) and the same as one launch per sweep (mixed)."""
    nz, nx, ny, h, (zc, xc, yc), nsweep = case
    vel = hetero(nz, nx, ny, h, 5)
    zs, xs, ys = zc * h, xc * h, yc * h
    ref, _, _ = oracle3d(vel, nz, nx, ny, h, zs, xs, ys, nsweep, 5.0)
    L = lib()
    ierr = c_int()
    g = GPU3D(nz, nx, ny, h, nsweep=nsweep, precision=precision)
    g.set_cells(vel)
    g.set_sources(zs, xs, ys)
    on, t, t32 = c_int(), c_int(), c_int()
    L.fteik_solver3d_setSweepOverlap(1, byref(ierr))   # small grid: chained
    L.fteik_solver3d_getSweepOverlap(byref(on), byref(t), byref(t32),
                                     byref(ierr))
    assert ierr.value == 0 and on.value == 1 and t.value >= 0
    L.fteik_solver3d_setSweepOverlap(0, byref(ierr))
    assert g.solve(1) == 0
    plain = g.field()
    if precision == 0:
        np.testing.assert_array_equal(plain, ref)
    L.fteik_solver3d_setSweepOverlap(2, byref(ierr))
    for tile in (10, 11, 12, 13):
        L.fteik_solver3d_setChainTile(tile, tile, byref(ierr))
        assert ierr.value == 0
        _announce("chain tile", tile, "precision", precision, "case", case)
        assert g.solve(1) == 0
        np.testing.assert_array_equal(g.field(), plain)


@pytest.mark.parametrize("precision", [0, 1])
def test_autotune_keeps_a_shape_that_matches_oracle(precision):
    """fteik_solver3d_autotune times every usable tile shape on the grid,
    keeps the fastest and leaves the solved field: it equals the oracle's
    (FP64), and a solve with the chosen shape gives the same field."""
    nz, nx, ny, h = 45, 52, 38, 15.0
    vel = hetero(nz, nx, ny, h, 18)
    zs, xs, ys = 30.2 * h, 8.4 * h, 20.7 * h
    ref, _, _ = oracle3d(vel, nz, nx, ny, h, zs, xs, ys, 3, 5.0)
    g = GPU3D(nz, nx, ny, h, nsweep=3, precision=precision)
    g.set_cells(vel)
    g.set_sources(zs, xs, ys)
    L = lib()
    t, t32, nt, ierr = c_int(-1), c_int(-1), c_int(), c_int(1)
    L.fteik_solver3d_autotune(1, byref(t), byref(t32), byref(ierr))
    assert ierr.value == 0
    _announce("autotuned tiles", t.value, t32.value, "precision", precision)
    cur, cur32 = c_int(), c_int()
    L.fteik_solver3d_getTile(byref(cur), byref(cur32), byref(nt), byref(ierr))
    assert (cur.value, cur32.value) == (t.value, t32.value)
    assert 0 <= t.value < nt.value
    tuned = g.field()
    if precision == 0:
        np.testing.assert_array_equal(tuned, ref)
    assert g.solve(1) == 0
    np.testing.assert_array_equal(g.field(), tuned)
