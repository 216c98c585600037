"""GPU parity of the 2D solver and the analytic fields (C-ABI) against the
reference's fixtures and the oracle."""
from ctypes import byref, c_int

import numpy as np
import pytest

from oracle import oracle as orc
from tests.gpu_helpers import lib, _d

pytestmark = pytest.mark.gpu


class GPU2D:
    def __init__(self, nz, nx, dz, dx, nsweep, eps, z0=0.0, x0=0.0):
        self.L = lib()
        self.nz, self.nx = nz, nx
        ierr = c_int(1)
        self.L.fteik_solver2d_initialize(nz, nx, z0, x0, dz, dx, nsweep, eps,
                                         0.0, 0, byref(ierr))
        assert ierr.value == 0

    def set_cells(self, vel, order=0):
        v = np.ascontiguousarray(vel, dtype=np.float64)
        ierr = c_int(1)
        self.L.fteik_solver2d_setCellVelocityModel64f(v.size, order, _d(v),
                                                      byref(ierr))
        assert ierr.value == 0

    def set_sources(self, zs, xs):
        z = np.ascontiguousarray(np.atleast_1d(zs), dtype=np.float64)
        x = np.ascontiguousarray(np.atleast_1d(xs), dtype=np.float64)
        ierr = c_int(1)
        self.L.fteik_solver2d_setSources64f(z.size, _d(z), _d(x), byref(ierr))
        assert ierr.value == 0
        self.nsrc = z.size

    def set_receivers(self, zr, xr):
        z = np.ascontiguousarray(zr, dtype=np.float64)
        x = np.ascontiguousarray(xr, dtype=np.float64)
        ierr = c_int(1)
        self.L.fteik_solver2d_setReceivers64f(z.size, _d(z), _d(x),
                                              byref(ierr))
        assert ierr.value == 0
        self.nrec = z.size

    def field(self, order=0):
        tt = np.zeros(self.nz * self.nx)
        ierr = c_int(1)
        self.L.fteik_solver2d_getTravelTimeField64f(tt.size, order, _d(tt),
                                                    byref(ierr))
        assert ierr.value == 0
        return tt

    def cells(self, order=0):
        tt = np.zeros((self.nz - 1) * (self.nx - 1))
        ierr = c_int(1)
        self.L.fteik_solver2d_getTravelTimeFieldCell64f(tt.size, order,
                                                        _d(tt), byref(ierr))
        assert ierr.value == 0
        return tt

    def receivers(self):
        ttr = np.zeros(self.nrec * self.nsrc)
        ierr = c_int(1)
        self.L.fteik_solver2d_getTravelTimes64f(self.nrec, _d(ttr),
                                                byref(ierr))
        assert ierr.value == 0
        return ttr.reshape(self.nsrc, self.nrec)


def _t2d_case(fixtures, case):
    nz, nx, dz, dx, zs, xs, nsweep, eps = fixtures["t2d_%s_cfg" % case]
    nz, nx, nsweep = int(nz), int(nx), int(nsweep)
    if case == "a":
        # testing/test2d.F90:40-46 velocity model
        z1 = (nz - 1) * dz
        iz = np.arange(1, nz)
        col = 3000.0 + (5000.0 - 3000.0) / z1 * ((iz - 1) * dz)
        vel = np.tile(col, nx - 1)
    else:
        vel = fixtures["t2d_b_vel"]
    return nz, nx, dz, dx, zs, xs, nsweep, eps, vel


@pytest.mark.parametrize("case", ["a", "b"])
def test_solve2d_matches_reference_fixtures(fixtures, case):
    """The reference's own 2D solves (test2d.F90 configuration and a
    heterogeneous model), bit for bit: field, receivers, cell averages."""
    nz, nx, dz, dx, zs, xs, nsweep, eps, vel = _t2d_case(fixtures, case)
    g = GPU2D(nz, nx, dz, dx, nsweep, eps)
    g.set_cells(vel)
    g.set_sources(zs, xs)
    zr, xr = fixtures["t2d_%s_rec" % case]
    g.set_receivers(zr, xr)
    ierr = c_int(1)
    g.L.fteik_solver2d_solveSourceLSM(1, byref(ierr))
    assert ierr.value == 0
    np.testing.assert_array_equal(g.field(), fixtures["t2d_%s_tt" % case])
    np.testing.assert_array_equal(g.receivers()[0],
                                  fixtures["t2d_%s_ttr" % case])
    if case == "b":
        np.testing.assert_array_equal(g.cells(1), fixtures["t2d_b_ttcell_xz"])
    # the FSM entry point against the reference's own FSM driver
    # (solver2D.F90:1042-1280, fixture from the reference build)
    g.L.fteik_solver2d_solveSourceFSM(1, byref(ierr))
    assert ierr.value == 0
    np.testing.assert_array_equal(g.field(), fixtures["t2d_%s_fsm_tt" % case])


def test_solve2d_many_sources_one_launch():
    """solveLSM over 9 sources (one workgroup each) equals per-source
    oracle solves; receiver table is [nrec x nsrc]."""
    nz, nx, h = 57, 83, 12.5
    rng = np.random.default_rng(77)
    from tests.golden.gen_golden_models import smooth_noise
    zc = (np.arange(nz - 1) + 0.5) * h
    vel = ((2000.0 + 1.1 * zc)[None, :] *
           (1.0 + 0.1 * smooth_noise(rng, (nx - 1, nz - 1), 3.0))).ravel()
    zs = rng.uniform(0.0, (nz - 2) * h, 9)
    xs = rng.uniform(0.0, (nx - 2) * h, 9)
    zr = np.zeros(7)
    xr = np.linspace(0.0, (nx - 1) * h, 7)
    g = GPU2D(nz, nx, h, h, 2, 3.0)
    g.set_cells(vel)
    g.set_sources(zs, xs)
    g.set_receivers(zr, xr)
    ierr = c_int(1)
    g.L.fteik_solver2d_solveLSM(byref(ierr))
    assert ierr.value == 0
    slow = orc.cell_slowness(vel, nz, nx)
    grid = (nz, nx, 1, h, h, 1.0, 0.0, 0.0, 0.0)
    ttr = g.receivers()
    for s in range(9):
        src = orc.source_locate(grid, zs[s], xs[s], slow=slow, is3d=False)
        ref = orc.solve2d(slow, nz, nx, h, h, src, 2, 3.0)
        np.testing.assert_array_equal(
            ttr[s], orc.receivers(ref, grid, zr, xr, is3d=False))
        if s == 8:       # the last source's field stays in the solver
            np.testing.assert_array_equal(g.field(), ref)


@pytest.mark.parametrize("nz,nx", [(57, 83), (45, 513), (33, 40),
                                   (120, 130), (64, 514)])
def test_solve2d_schedules_match_oracle(nz, nx):
    """Pipelined columns (one lane per x column, several waves for wide
    grids; grids wider than 513 nodes fall back to the level schedule) and
    the level-by-level schedule give the oracle's fields bit for bit, for
    sources near every corner and in the middle (initialization boxes of
    every extent)."""
    h = 10.0
    rng = np.random.default_rng(nz * 1000 + nx)
    from tests.golden.gen_golden_models import smooth_noise
    zc = (np.arange(nz - 1) + 0.5) * h
    vel = ((1800.0 + 0.9 * zc)[None, :] *
           (1.0 + 0.1 * smooth_noise(rng, (nx - 1, nz - 1), 3.0))).ravel()
    zs = np.array([0.2, nz - 2.1, 0.0, nz - 2.5, 0.5 * nz]) * h
    xs = np.array([0.3, 0.0, nx - 2.2, nx - 2.4, 0.5 * nx]) * h
    zr = np.zeros(5)
    xr = np.linspace(0.0, (nx - 1) * h, 5)
    slow = orc.cell_slowness(vel, nz, nx)
    grid = (nz, nx, 1, h, h, 1.0, 0.0, 0.0, 0.0)
    refs = []
    for k in range(zs.size):
        src = orc.source_locate(grid, zs[k], xs[k], slow=slow, is3d=False)
        refs.append(orc.solve2d(slow, nz, nx, h, h, src, 2, 3.0))
    ierr = c_int(1)
    for sched in (0, 1):
        print("case: 2D", nz, nx, "schedule", sched, flush=True)
        g = GPU2D(nz, nx, h, h, 2, 3.0)
        g.L.fteik_solver2d_setSchedule(sched, byref(ierr))
        assert ierr.value == 0
        g.set_cells(vel)
        g.set_receivers(zr, xr)
        for k in range(zs.size):
            g.set_sources(zs[k], xs[k])
            g.L.fteik_solver2d_solveSourceLSM(1, byref(ierr))
            assert ierr.value == 0
            np.testing.assert_array_equal(g.field(), refs[k])
        # all sources in one launch: receiver times
        g.set_sources(zs, xs)
        g.L.fteik_solver2d_solveLSM(byref(ierr))
        assert ierr.value == 0
        ttr = g.receivers()
        for k in range(zs.size):
            np.testing.assert_array_equal(
                ttr[k], orc.receivers(refs[k], grid, zr, xr, is3d=False))
    g.L.fteik_solver2d_setSchedule(2, byref(ierr))
    assert ierr.value == 1


def test_solve2d_error_paths():
    L = lib()
    g = GPU2D(10, 12, 10.0, 10.0, 2, 3.0)
    ierr = c_int(0)
    L.fteik_solver2d_solveLSM(byref(ierr))
    assert ierr.value == 1               # no model
    g.set_cells(np.full(9 * 11, 2500.0))
    g.set_sources(85.0, 30.0)            # zsi = nz-1: last cell
    L.fteik_solver2d_solveSourceLSM(1, byref(ierr))
    assert ierr.value == 1
    L.fteik_solver2d_solveSourceLSM(3, byref(ierr))
    assert ierr.value == 1


# -------------------------------------------------------------- analytic
def test_analytic_fields_match_reference(fixtures):
    nz, nx, ny, dz, dx, dy, zs, xs, ys, v0, v1 = fixtures["ana_cfg"]
    nz, nx, ny = int(nz), int(nx), int(ny)
    L = lib()
    ierr = c_int(1)
    L.fteik_analytic_initialize64f(nz, nx, ny, 0.0, 0.0, 0.0, dz, dx, dy, 0,
                                   byref(ierr))
    assert ierr.value == 0
    z, x, y = (np.array([v]) for v in (zs, xs, ys))
    L.fteik_analytic_setSources64f(1, _d(z), _d(x), _d(y), byref(ierr))
    assert ierr.value == 0
    n = nz * nx * ny
    L.fteik_analytic_setLinearVelocityGradient64f(v0, v1, byref(ierr))
    assert ierr.value == 0
    L.fteik_analytic_solveSourceLinearVelocityGradient64f(1, byref(ierr))
    assert ierr.value == 0
    # device acosh/log/sqrt are not correctly rounded: 1e-13 relative
    for order, key, tol in ((0, "ana_grad_natural", 1e-13),
                            (1, "ana_grad_xyz", 1e-13),
                            (2, "ana_grad_zyx", 1e-13)):
        tt = np.zeros(n)
        L.fteik_analytic_getTravelTimeField64f(n, order, _d(tt), byref(ierr))
        assert ierr.value == 0
        np.testing.assert_allclose(tt, fixtures[key], rtol=tol, atol=0)
    L.fteik_analytic_setConstantVelocity64f(4200.0, byref(ierr))
    L.fteik_analytic_solveSourceConstantVelocity64f(1, byref(ierr))
    assert ierr.value == 0
    tt = np.zeros(n)
    L.fteik_analytic_getTravelTimeField64f(n, 0, _d(tt), byref(ierr))
    np.testing.assert_allclose(tt, fixtures["ana_const_natural"], rtol=1e-15,
                               atol=0)
    # receivers: the closed form at arbitrary points
    zr = np.array([0.0, 13.0, 40.0])
    xr = np.array([5.0, 77.0, 21.5])
    yr = np.array([1.0, 9.0, 30.0])
    L.fteik_analytic_setReceivers64f(3, _d(zr), _d(xr), _d(yr), byref(ierr))
    assert ierr.value == 0
    trec = np.zeros(3)
    L.fteik_analytic_getTravelTimesConstantVel64f(3, _d(trec), byref(ierr))
    assert ierr.value == 0
    r = np.sqrt((zr - zs) ** 2 + (xr - xs) ** 2 + (yr - ys) ** 2)
    np.testing.assert_allclose(trec, r / 4200.0, rtol=1e-15)
    L.fteik_analytic_free()
