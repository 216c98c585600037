"""The oracle's travel-time gradients and ray tracer, pinned on the CPU.

- 2D gradient (solver2D.F90:984-1032): bit for bit against the reference
  build's own gradient of its solved fields (fixtures).
- 3D gradient (solver3D.F90:1249-1321; the reference's 3D solver does not
  run from its build): forward differences of a quadratic are the
  derivative at the midpoint, exactly, and backward ones at the last face;
  distinct coefficients per axis catch an exchanged axis, a wrong sign or
  an off-by-one index.
- Rays (the reference documents the method, rays.F90:846-863, but its
  tracers do not run): in a homogeneous medium they run straight to the
  source and their time is distance / velocity; in a linear gradient their
  time approaches the closed-form travel time (analytic.F90:456-528).
"""
import numpy as np
import pytest

from oracle import oracle as orc


def _coords(nz, nx, ny, h):
    iz, ix, iy = np.meshgrid(np.arange(nz), np.arange(nx), np.arange(ny),
                             indexing="ij")
    f = lambda a: (a * h).transpose(2, 1, 0).ravel()   # natural order
    return f(iz), f(ix), f(iy)


@pytest.mark.parametrize("case", ["a", "b"])
def test_gradient2d_matches_reference(fixtures, case):
    nz, nx, dz, dx = (fixtures["t2d_%s_cfg" % case][i] for i in range(4))
    nz, nx = int(nz), int(nx)
    gx, gz = orc.fd_gradient2d(fixtures["t2d_%s_tt" % case], nz, nx, dz, dx)
    np.testing.assert_array_equal(gx, fixtures["t2d_%s_gradx" % case])
    np.testing.assert_array_equal(gz, fixtures["t2d_%s_gradz" % case])


def test_gradient3d_quadratic_is_exact():
    nz, nx, ny, h = 7, 9, 6, 1.0
    Z, X, Y = _coords(nz, nx, ny, h)
    a, b, c, d = 3.0, 5.0, 7.0, 2.0
    tt = a * Z ** 2 + b * X ** 2 + c * Y ** 2 + d * Z * X
    g = orc.fd_gradient3d(tt, nz, nx, ny, h, h, h).reshape(-1, 3)
    # forward difference = derivative at the midpoint; the last node of an
    # axis takes the backward difference (the midpoint below it)
    zm = np.where(Z < (nz - 1) * h, Z + 0.5 * h, Z - 0.5 * h)
    xm = np.where(X < (nx - 1) * h, X + 0.5 * h, X - 0.5 * h)
    ym = np.where(Y < (ny - 1) * h, Y + 0.5 * h, Y - 0.5 * h)
    np.testing.assert_array_equal(g[:, 0], 2 * b * xm + d * Z)
    np.testing.assert_array_equal(g[:, 1], 2 * c * ym)
    np.testing.assert_array_equal(g[:, 2], 2 * a * zm + d * X)


def test_gradient2d_quadratic_is_exact():
    nz, nx, h = 8, 11, 1.0
    iz, ix = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    Z = (iz * h).T.ravel()
    X = (ix * h).T.ravel()
    tt = 3.0 * Z ** 2 + 5.0 * X ** 2
    gx, gz = orc.fd_gradient2d(tt, nz, nx, h, h)
    # backward differences, forward on the first row / column
    zm = np.where(Z > 0, Z - 0.5 * h, Z + 0.5 * h)
    xm = np.where(X > 0, X - 0.5 * h, X + 0.5 * h)
    np.testing.assert_array_equal(gz, 6.0 * zm)
    np.testing.assert_array_equal(gx, 10.0 * xm)


def _homogeneous(nz, nx, ny, h, v, src):
    Z, X, Y = _coords(nz, nx, ny, h)
    zs, xs, ys = (i * h for i in src)
    tt = np.sqrt((Z - zs) ** 2 + (X - xs) ** 2 + (Y - ys) ** 2) / v
    slow = np.full((nz - 1) * (nx - 1) * (ny - 1), 1.0 / v)
    return tt, slow, (xs, ys, zs)


@pytest.mark.parametrize("integrator", [1, 2, 3])
def test_rays_straight_in_homogeneous_medium(integrator):
    nz, nx, ny, h, v = 31, 35, 29, 10.0, 3000.0
    tt, slow, s = _homogeneous(nz, nx, ny, h, v, (10, 12, 9))
    g = orc.fd_gradient3d(tt, nz, nx, ny, h, h, h)
    rng = np.random.default_rng(integrator)
    for _ in range(12):
        rec = (rng.uniform(0, (nx - 1) * h), rng.uniform(0, (ny - 1) * h),
               rng.uniform(0, (nz - 1) * h))
        ok, pts, t = orc.trace_ray3d(slow, g, nz, nx, ny, h, h, h, integrator,
                                     s, rec, 4 * (nx + ny + nz) + 16)
        assert ok
        a, b = np.array(rec), np.array(s)
        np.testing.assert_array_equal(pts[0], a)
        np.testing.assert_array_equal(pts[-1], b)
        dist = np.linalg.norm(b - a)
        if dist < 2 * h:
            continue
        u = (b - a) / dist
        off = np.linalg.norm((pts - a) - np.outer((pts - a) @ u, u), axis=1)
        # first-order gradients sit half a cell off the node: rays stay
        # within a cell of the straight line, and their time within 1 %
        # plus a quarter cell (the last step snaps onto the source node)
        assert off.max() < h
        assert abs(t - dist / v) <= 0.01 * dist / v + 0.25 * h / v


def test_ray_from_the_source_ends_at_once():
    nz, nx, ny, h, v = 12, 13, 11, 10.0, 2500.0
    tt, slow, s = _homogeneous(nz, nx, ny, h, v, (5, 6, 4))
    g = orc.fd_gradient3d(tt, nz, nx, ny, h, h, h)
    ok, pts, t = orc.trace_ray3d(slow, g, nz, nx, ny, h, h, h, 3, s, s, 100)
    assert ok and pts.shape == (2, 3)
    np.testing.assert_array_equal(pts[1], s)
    assert t == 0.0


@pytest.mark.parametrize("integrator", [1, 3])
def test_rays_follow_linear_gradient_closed_form(integrator):
    """v(z) = 3000 + 0.8 z: the ray time from a surface receiver to a deep
    source is within 1 % of the closed-form travel time."""
    nz, nx, ny, h = 41, 45, 39, 25.0
    vtop = 3000.0
    vbot = vtop + 0.8 * (nz - 1) * h
    zs, xs, ys = 30 * h, 22 * h, 19 * h
    tt = orc.xfast_to_natural(
        orc.analytic_gradient(nz, nx, ny, h, h, h, 0.0, 0.0, 0.0, zs, xs, ys,
                              vtop, vbot), nz, nx, ny)
    zc = (np.arange(nz - 1) + 0.5) * h
    slow = np.tile(1.0 / (vtop + 0.8 * zc), (nx - 1) * (ny - 1))
    g = orc.fd_gradient3d(tt, nz, nx, ny, h, h, h)
    for rec in ((0.0, 0.0, 0.0), (40 * h, 5 * h, 0.0), (10 * h, 38 * h, 0.0),
                (44 * h, 38 * h, 12 * h)):
        ok, pts, t = orc.trace_ray3d(slow, g, nz, nx, ny, h, h, h, integrator,
                                     (xs, ys, zs), rec, 4 * (nx + ny + nz))
        assert ok
        ta = orc.analytic_gradient_point(nz, h, 0.0, zs, xs, ys, vtop, vbot,
                                         rec[2], rec[0], rec[1])
        assert abs(t - ta) <= 0.01 * ta, (rec, t, ta)


def test_rays2d_homogeneous():
    nz, nx, h, v = 30, 40, 15.0, 2000.0
    iz, ix = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    Z = (iz * h).T.ravel()
    X = (ix * h).T.ravel()
    zs, xs = 11 * h, 27 * h
    tt = np.sqrt((Z - zs) ** 2 + (X - xs) ** 2) / v
    gx, gz = orc.fd_gradient2d(tt, nz, nx, h, h)
    slow = np.full((nz - 1) * (nx - 1), 1.0 / v)
    for integrator in (1, 2, 3):
        for rec in ((0.0, 0.0), (39 * h, 0.0), (2 * h, 29 * h)):
            ok, pts, t = orc.trace_ray2d(slow, gx, gz, nz, nx, h, h,
                                         integrator, (xs, zs), rec, 400)
            assert ok
            assert np.all(pts[:, 1] == 0.0)
            d = np.hypot(rec[0] - xs, rec[1] - zs)
            assert abs(t - d / v) <= 0.01 * d / v
