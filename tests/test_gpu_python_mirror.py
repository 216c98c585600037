"""The Python mirror classes of the reference's bindings (python/fteik3d.py,
python/fteik2d.py, python/fteikAnalytic.py) driven end to end on the GPU, with
the reference bindings' array conventions: models given as [nz-1, ny-1, nx-1]
(3D) or [nz-1, nx-1] (2D) cell arrays flattened in Fortran order with the
ZYX / ZX ordering, fields returned in the same shapes.  Results equal the
oracle's natural-order fields after the same reshapes, bit for bit (FP64
solvers) or to 1e-13 (closed forms)."""
import numpy as np
import pytest

import fteik_amd as fa
from oracle import oracle as orc
from tests.gpu_helpers import hetero

pytestmark = pytest.mark.gpu


def test_fteik3d_mirror_end_to_end():
    nz, nx, ny, h = 17, 21, 15, 30.0
    vnat = hetero(nz, nx, ny, h, 71)               # natural: z fastest
    V = vnat.reshape(ny - 1, nx - 1, nz - 1).transpose(2, 0, 1)  # [z, y, x]
    xs, ys, zs = np.array([5.5 * h, 14.2 * h]), np.array([3.3 * h, 9.1 * h]), \
        np.array([7.7 * h, 1.0 * h])
    xr = np.linspace(0.0, (nx - 1) * h, 6)
    yr = np.linspace(0.0, (ny - 1) * h, 6)[::-1].copy()
    zr = np.zeros(6)
    with fa.fteik3d() as s:
        assert s.initialize(nx, ny, nz, h, h, h, nsweep=2, eps=5.0) == 0
        assert s.setVelocityModel(V) == 0
        assert s.setSources(xs, ys, zs) == 0
        assert s.setReceivers(xr, yr, zr) == 0
        slow = orc.cell_slowness(vnat, nz, nx, ny)
        grid = (nz, nx, ny, h, h, h, 0.0, 0.0, 0.0)
        for k in range(2):
            assert s.solveSource(k + 1) == 0
            src = orc.source_locate(grid, zs[k], xs[k], ys[k], slow)
            ref = orc.solve3d(slow, nz, nx, ny, h, h, h, src, 2, 5.0,
                              schedule=1)
            tt = s.getTravelTimeField()
            assert tt.shape == (nz, ny, nx)
            np.testing.assert_array_equal(
                tt, ref.reshape(ny, nx, nz).transpose(2, 0, 1))
            np.testing.assert_array_equal(
                s.getTravelTimes(), orc.receivers(ref, grid, zr, xr, yr))
            cell = s.getTravelTimeField(wantCell=True)
            assert cell.shape == (nz - 1, ny - 1, nx - 1)
            np.testing.assert_array_equal(
                np.reshape(cell, cell.size, order="F"),
                orc.field_cell3d(ref, nz, nx, ny, 2))
        assert s.solveLSM() == 0          # every source in turn


def test_fteik2d_mirror_end_to_end():
    nz, nx, h = 31, 44, 20.0
    rng = np.random.default_rng(8)
    zc = (np.arange(nz - 1) + 0.5) * h
    V = (2200.0 + 0.9 * zc)[:, None] * (1.0 + 0.05 * rng.random((nz - 1,
                                                                nx - 1)))
    vnat = np.reshape(V, V.size, order="F")        # [z, x] -> z fastest
    xs = np.array([3.1 * h, 30.4 * h, 12.0 * h])
    zs = np.array([2.2 * h, 20.5 * h, 0.0])
    xr = np.linspace(0.0, (nx - 1) * h, 5)
    zr = np.zeros(5)
    with fa.fteik2d() as s:
        assert s.initialize(nx, nz, h, h, nsweep=2, eps=3.0) == 0
        assert s.setVelocityModel(V) == 0
        assert s.setSources(xs, zs) == 0
        assert s.setReceivers(xr, zr) == 0
        assert s.solveLSM() == 0
        ttr = s.getTravelTimes()
        assert ttr.shape == (5, 3)
        slow = orc.cell_slowness(vnat, nz, nx)
        grid = (nz, nx, 1, h, h, 1.0, 0.0, 0.0, 0.0)
        for k in range(3):
            src = orc.source_locate(grid, zs[k], xs[k], slow=slow, is3d=False)
            ref = orc.solve2d(slow, nz, nx, h, h, src, 2, 3.0)
            np.testing.assert_array_equal(
                ttr[:, k], orc.receivers(ref, grid, zr, xr, is3d=False))
        tt = s.getTravelTimeField()                # last source's field
        assert tt.shape == (nz, nx)
        np.testing.assert_array_equal(np.reshape(tt, tt.size, order="F"), ref)
        assert s.solveSource(2, fsm=True) == 0
        src = orc.source_locate(grid, zs[1], xs[1], slow=slow, is3d=False)
        ref = orc.solve2d(slow, nz, nx, h, h, src, 2, 3.0)
        np.testing.assert_array_equal(
            np.reshape(s.getTravelTimeField(), nz * nx, order="F"), ref)


def test_fteik_analytic_mirror():
    nz, nx, ny, h = 12, 15, 9, 50.0
    xs, ys, zs = 3.3 * h, 4.4 * h, 5.5 * h
    a = fa.fteikAnalytic()
    assert a.initialize(nx, ny, nz, h, h, h) == 0
    assert a.setSources([xs], [ys], [zs]) == 0
    assert a.setLinearVelocityGradient(3000.0, 4500.0) == 0
    assert a.solveLinearVelocityGradient() == 0
    tt = a.getTravelTimeField()
    assert tt.shape == (nz, ny, nx)
    nat = orc.xfast_to_natural(
        orc.analytic_gradient(nz, nx, ny, h, h, h, 0.0, 0.0, 0.0, zs, xs, ys,
                              3000.0, 4500.0), nz, nx, ny)
    np.testing.assert_allclose(tt, nat.reshape(ny, nx, nz).transpose(2, 0, 1),
                               rtol=1e-13, atol=0)
    assert a.setReceivers([0.0, 200.0], [10.0, 300.0], [0.0, 100.0]) == 0
    assert a.setConstantVelocity(4000.0) == 0
    trc = a.getTravelTimesConstantVelocity()
    r = np.sqrt((np.array([0.0, 200.0]) - xs) ** 2 +
                (np.array([10.0, 300.0]) - ys) ** 2 +
                (np.array([0.0, 100.0]) - zs) ** 2)
    np.testing.assert_allclose(trc[:, 0], r / 4000.0, rtol=1e-15)
    a.free()


def test_locator_class():
    nz, nx, ny, h = 9, 11, 8, 100.0
    ngrd = nz * nx * ny
    rng = np.random.default_rng(4)
    tabs = rng.uniform(0.5, 3.0, (4, ngrd))
    loc = fa.Locator(3, 4, ngrd)
    try:
        loc.setReferenceCompatibility(0)
        for k in range(4):
            loc.setTravelTimeField(k + 1, tabs[k] if k % 2 else
                                   tabs[k].astype(np.float32))
        ref = orc.Locator(3, 4, ngrd)
        for k in range(4):
            ref.set_table(k + 1, tabs[k])
        for e in range(3):
            cell = int(rng.integers(0, ngrd))
            picks = tabs[:, cell] + 7.0
            loc.setObservation(e + 1, [1, 2, 3, 4], picks)
            ref.set_observation(e + 1, np.arange(1, 5), picks, np.ones(4))
            want = ref.locate(e + 1, accumulate_t0=False, pad_argmin=False)
            assert loc.locateEvent(e + 1) == want
            assert want[0] == cell + 1
        opt, t0, obj = loc.locateEvents([3, 1])
        assert opt[0] == ref.locate(3, False, False)[0]
    finally:
        loc.setReferenceCompatibility(1)
        loc.finalize()
