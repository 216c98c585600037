"""GPU travel-time gradients and ray tracing through the C-ABI against the
oracle, bit for bit: the 3D gradient of a solved heterogeneous field, the
2D gradient against the reference build's own fixtures, and rays (paths,
point counts, times) for every integrator, traced to points and to the
receivers."""
import ctypes
from ctypes import byref, c_double, c_int

import numpy as np
import pytest

from oracle import oracle as orc
from tests.gpu_helpers import GPU3D, hetero, lib, oracle3d, _d

pytestmark = pytest.mark.gpu


def _rays(L, n):
    nr, ierr = c_int(), c_int()
    L.fteik_rays_getNumberOfRays(byref(nr), byref(ierr))
    assert ierr.value == 0 and nr.value == n
    times = np.zeros(n)
    L.fteik_rays_getTravelTimes64f(n, _d(times), byref(ierr))
    assert ierr.value == 0
    paths = []
    for r in range(n):
        xyz = np.zeros(3 * 20000)
        npts, t = c_int(), c_double()
        L.fteik_rays_getRayPath(r + 1, 20000, _d(xyz), byref(npts), byref(t),
                                byref(ierr))
        assert ierr.value == 0 and t.value == times[r]
        paths.append(xyz[:3 * npts.value].reshape(npts.value, 3))
    return paths, times


def test_rays3d_match_oracle():
    nz, nx, ny, h = 26, 31, 23, 20.0
    origin = (50.0, -100.0, 25.0)
    vel = hetero(nz, nx, ny, h, 66)
    zs, xs, ys = origin[0] + 17.4 * h, origin[1] + 9.2 * h, \
        origin[2] + 11.6 * h
    ref, slow, src = oracle3d(vel, nz, nx, ny, h, zs, xs, ys, 3, 5.0,
                              origin=origin)
    g = GPU3D(nz, nx, ny, h, nsweep=3, origin=origin)
    g.set_cells(vel)
    g.set_sources(zs, xs, ys)
    rng = np.random.default_rng(6)
    nr = 9
    xr = origin[1] + rng.uniform(0, (nx - 1) * h, nr)
    yr = origin[2] + rng.uniform(0, (ny - 1) * h, nr)
    zr = origin[0] + np.concatenate([np.zeros(5),
                                     rng.uniform(0, (nz - 1) * h, 4)])
    g.set_receivers(zr, xr, yr)
    assert g.solve(1) == 0
    np.testing.assert_array_equal(g.field(), ref)
    L = lib()
    ierr = c_int(1)
    n = nz * nx * ny
    grad = np.zeros(3 * n)
    L.fteik_solver3d_getTravelTimeGradient64f(n, _d(grad), byref(ierr))
    assert ierr.value == 0
    want = orc.fd_gradient3d(ref, nz, nx, ny, h, h, h)
    np.testing.assert_array_equal(grad, want)
    # the rays end at the true source position (grid-relative)
    snode = (xs - origin[1], ys - origin[2], zs - origin[0])
    maxpts = 4 * (nx + ny + nz) + 16
    for integ in (1, 2, 3):
        print("case: rays3d integrator", integ, flush=True)
        L.fteik_rays_setODEIntegrator(integ, byref(ierr))
        assert ierr.value == 0
        zp, xp, yp = (np.ascontiguousarray(a) for a in (zr, xr, yr))
        L.fteik_solver3d_traceRaysToPoints(nr, _d(zp), _d(xp), _d(yp),
                                           byref(ierr))
        assert ierr.value == 0
        paths, times = _rays(L, nr)
        for r in range(nr):
            rec = (xr[r] - origin[1], yr[r] - origin[2], zr[r] - origin[0])
            ok, pts, t = orc.trace_ray3d(slow, want, nz, nx, ny, h, h, h,
                                         integ, snode, rec, maxpts)
            assert ok
            assert times[r] == t
            np.testing.assert_array_equal(
                paths[r], pts + np.array([origin[1], origin[2], origin[0]]))
        # to the receivers: the same rays
        L.fteik_solver3d_traceRaysToReceivers(byref(ierr))
        assert ierr.value == 0
        p2, t2 = _rays(L, nr)
        np.testing.assert_array_equal(t2, times)
        # a ray's time is the field's time at its receiver, to first order
        ttr = g.receivers()
        assert np.all(np.abs(times - ttr) <= 0.03 * ttr + 0.5 * h / 1500.0)
    L.fteik_rays_setODEIntegrator(1, byref(ierr))


def test_rays3d_error_paths():
    L = lib()
    g = GPU3D(12, 11, 10, 10.0)
    ierr = c_int(0)
    z = np.array([5.0])
    L.fteik_solver3d_traceRaysToPoints(1, _d(z), _d(z), _d(z), byref(ierr))
    assert ierr.value == 1                   # no solved field
    g.set_cells(np.full(11 * 10 * 9, 3000.0))
    g.set_sources(50.0, 50.0, 50.0)
    assert g.solve(1) == 0
    far = np.array([1e6])
    L.fteik_solver3d_traceRaysToPoints(1, _d(z), _d(far), _d(z), byref(ierr))
    assert ierr.value == 1                   # outside the model
    L.fteik_rays_setODEIntegrator(7, byref(ierr))
    assert ierr.value == 1
    L.fteik_solver3d_traceRaysToReceivers(byref(ierr))
    assert ierr.value == 1                   # no receivers
    L.fteik_solver3d_traceRaysToPoints(1, _d(z), _d(z), _d(z), byref(ierr))
    assert ierr.value == 0                   # Euler after the bad value


@pytest.mark.parametrize("case", ["a", "b"])
def test_rays2d_gradient_matches_reference_and_rays_match_oracle(fixtures,
                                                                 case):
    from tests.test_gpu_solver2d import GPU2D, _t2d_case
    nz, nx, dz, dx, zs, xs, nsweep, eps, vel = _t2d_case(fixtures, case)
    g = GPU2D(nz, nx, dz, dx, nsweep, eps)
    g.set_cells(vel)
    g.set_sources(zs, xs)
    zr = np.array([0.0, 0.3 * (nz - 1) * dz, (nz - 1) * dz])
    xr = np.array([0.0, 0.9 * (nx - 1) * dx, 0.5 * (nx - 1) * dx])
    g.set_receivers(zr, xr)
    L = lib()
    ierr = c_int(1)
    L.fteik_solver2d_solveSourceLSM(1, byref(ierr))
    assert ierr.value == 0
    np.testing.assert_array_equal(g.field(), fixtures["t2d_%s_tt" % case])
    n = nz * nx
    gx, gz = np.zeros(n), np.zeros(n)
    L.fteik_solver2d_getTravelTimeGradient64f(n, _d(gx), _d(gz), byref(ierr))
    assert ierr.value == 0
    np.testing.assert_array_equal(gx, fixtures["t2d_%s_gradx" % case])
    np.testing.assert_array_equal(gz, fixtures["t2d_%s_gradz" % case])
    slow = orc.cell_slowness(vel, nz, nx)
    snode = (xs, zs)
    for integ in (1, 3):
        L.fteik_rays_setODEIntegrator(integ, byref(ierr))
        L.fteik_solver2d_traceRaysToReceivers(byref(ierr))
        assert ierr.value == 0
        paths, times = _rays(L, 3)
        for r in range(3):
            ok, pts, t = orc.trace_ray2d(slow, gx, gz, nz, nx, dz, dx, integ,
                                         snode, (xr[r], zr[r]),
                                         4 * (nx + 1 + nz) + 16)
            assert ok and times[r] == t
            np.testing.assert_array_equal(paths[r], pts)
        # a ray's time is the field's time at its receiver, to first order
        ttr = g.receivers()[0]
        assert np.all(np.abs(times - ttr) <= 0.02 * ttr)
    L.fteik_rays_setODEIntegrator(1, byref(ierr))
