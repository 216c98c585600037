"""Pins the CPU oracle (oracle/fteik_oracle.c) against the reference.

Bitwise checks against the fixtures written by tests/golden/gen_golden.py
from the reference build, plus the properties the mathematics fixes for the
parts the reference build cannot run (its 3D level-set driver): level
structure by brute force, level-order invariance, closed-form solutions and
the original Noble-Gesret 3D solver.  CPU only.
"""
import numpy as np
import pytest

from oracle import oracle as orc

HUGE = orc.HUGE


def grid9(nz, nx, ny, h, z0=0.0, x0=0.0, y0=0.0):
    return (nz, nx, ny, h, h, h, z0, x0, y0)


# ------------------------------------------------------------------ model
def test_index_round_trip():
    """testing/tests.c:262-302 on the same 34 x 10 x 21 grid."""
    import ctypes
    L = orc.lib()
    nz, nx, ny = 34, 10, 21
    i1, j1, k1 = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    for k in range(1, ny + 1):
        for j in range(1, nx + 1):
            for i in range(1, nz + 1):
                g = L.or_grid2index(i, j, k, nz, nz * nx)
                assert g == (k - 1) * nz * nx + (j - 1) * nz + i
                ierr = L.or_index2grid(g, nz, nx, ny, ctypes.byref(i1),
                                       ctypes.byref(j1), ctypes.byref(k1))
                assert ierr == 0 and (i1.value, j1.value, k1.value) == (i, j, k)


@pytest.mark.parametrize("kind,order", [("cell", 1), ("cell", 2),
                                        ("nodal", 0), ("nodal", 1),
                                        ("nodal", 2)])
def test_model_orders(fixtures, kind, order):
    nz, nx, ny = (int(v) for v in fixtures["model_grid"])
    vin = fixtures["model_%s%d_in" % (kind, order)]
    if kind == "cell":
        s = orc.cell_slowness(vin, nz, nx, ny, order)
    else:
        s = orc.nodal_slowness(vin, nz, nx, ny, order)
    np.testing.assert_array_equal(1.0 / s,
                                  fixtures["model_%s%d_velback" % (kind,
                                                                   order)])


def test_source_locate(fixtures):
    g = fixtures["source_grid"]
    grid = (int(g[0]), int(g[1]), int(g[2])) + tuple(g[3:])
    nz, nx, ny = grid[:3]
    rng = np.random.default_rng(20240601)
    # the fixture's source cases used the velocity model of gen_golden's
    # first block; rebuild it to get the same szero
    from tests.golden.gen_golden_models import hetero_model3d
    vel = hetero_model3d(rng, nz, nx, ny, grid[3])
    slow = 1.0 / vel
    for s, ints, reals in zip(fixtures["source_in"], fixtures["source_ints"],
                              fixtures["source_reals"]):
        src = orc.source_locate(grid, *s, slow=slow)
        assert (src.zsi, src.xsi, src.ysi) == tuple(ints)
        assert (src.zsa, src.xsa, src.ysa, src.szero) == tuple(reals)


def test_receivers(fixtures):
    nz, nx, ny = (int(v) for v in fixtures["model_grid"])
    z, x, y = fixtures["rec_zxy"]
    ttr = orc.receivers(fixtures["rec_field"], grid9(nz, nx, ny, 25.0), z, x,
                        y)
    np.testing.assert_array_equal(ttr, fixtures["rec_ttr"])


def test_gathers(fixtures):
    nz, nx, ny = (int(v) for v in fixtures["model_grid"])
    slow = fixtures["gather_slow"]
    tt = fixtures["gather_tt"]
    for (sweep, i, j, k), vals in zip(fixtures["gather_nodes"],
                                      fixtures["gather_vals"]):
        s7 = orc.gather_slow3d(int(sweep), int(i), int(j), int(k), nz, nx, ny,
                               slow)
        t7, t8 = orc.gather_tt3d(int(sweep), int(i), int(j), int(k), nz, nx,
                                 tt)
        np.testing.assert_array_equal(s7, vals[:7])
        np.testing.assert_array_equal(t7, vals[7:14])
        assert t8 == vals[14]


# ------------------------------------------------------ 3D local solvers
def _ls3d_consts(fixtures):
    nz, nx, ny, h, zs, xs, ys, eps = fixtures["ls3d_grid"]
    nz, nx, ny = int(nz), int(nx), int(ny)
    slow = 1.0 / fixtures["ls3d_vel"]
    src = orc.source_locate(grid9(nz, nx, ny, h), zs, xs, ys, slow)
    return orc.consts3d(h, h, h, eps, src), (nz, nx, ny)


def test_seeds3d(fixtures):
    import ctypes
    c, (nz, nx, ny) = _ls3d_consts(fixtures)
    dest = np.zeros(8, dtype=np.int32)
    ts8 = np.zeros(8)
    orc.lib().or_seeds3d(ctypes.byref(c), nz, nx, orc._i(dest), orc._d(ts8))
    np.testing.assert_array_equal(ts8, fixtures["ls3d_ts8"])
    np.testing.assert_array_equal(dest + 1, fixtures["ls3d_dest"])


def test_local_solvers3d(fixtures):
    import ctypes
    c, _ = _ls3d_consts(fixtures)
    L = orc.lib()
    rows = fixtures["ls3d_in"]
    nodes = fixtures["ls3d_node"]
    got_n, got_i = [], []
    for r, (sweep, i, j, k) in zip(rows, nodes):
        t7 = np.ascontiguousarray(r[:7])
        s7 = np.ascontiguousarray(r[8:15])
        got_n.append(L.or_noinit3d(ctypes.byref(c), orc._d(t7), orc._d(s7),
                                   r[7]))
        got_i.append(L.or_init3d(ctypes.byref(c), orc._d(t7), orc._d(s7),
                                 r[7], int(i), int(j), int(k), int(sweep)))
    np.testing.assert_array_equal(np.array(got_n), fixtures["ls3d_noinit"])
    np.testing.assert_array_equal(np.array(got_i), fixtures["ls3d_init"])
    # the fixtures must exercise the interesting branches
    ref_i = fixtures["ls3d_init"]
    assert np.sum(ref_i < rows[:, 7]) > 100     # init improves the node
    assert np.sum(fixtures["ls3d_noinit"] < rows[:, 7]) > 100


# ------------------------------------------------------------------- 2D
def test_local_solvers2d(fixtures):
    nz, nx, dz, dx, zs, xs, eps = fixtures["ls2d_cfg"]
    nz, nx = int(nz), int(nx)
    slow = 1.0 / fixtures["ls2d_vel"]
    src = orc.source_locate((nz, nx, 1, dz, dx, 1.0, 0.0, 0.0, 0.0), zs, xs,
                            slow=slow, is3d=False)
    L = orc.lib()
    import ctypes
    for row in fixtures["ls2d_rows"]:
        sweep, iz, ix = (int(v) for v in row[:3])
        sgntz = 1 if sweep <= 2 else -1
        sgntx = 1 if sweep in (1, 3) else -1
        tt = np.full(nz * nx, HUGE)
        tv, te, tev, t0 = row[5:9]
        tt[(ix - 1) * nz + iz - sgntz - 1] = tv
        tt[(ix - sgntx - 1) * nz + iz - 1] = te
        tt[(ix - sgntx - 1) * nz + iz - sgntz - 1] = tev
        tt[(ix - 1) * nz + iz - 1] = t0
        got = L.or_init2d(sweep, nz, nx, iz, ix, dz, dx, eps,
                          ctypes.byref(src), orc._d(slow), orc._d(tt))
        assert got == row[3]
        t4 = np.ascontiguousarray(row[5:9])
        s3 = np.ascontiguousarray(row[9:12])
        assert L.or_noinit2d(orc._d(t4), orc._d(s3), dz, dx) == row[4]


@pytest.mark.parametrize("case", ["a", "b"])
def test_solve2d_matches_reference(fixtures, case):
    cfg = fixtures["t2d_%s_cfg" % case]
    nz, nx, dz, dx, zs, xs, nsweep, eps = cfg
    nz, nx, nsweep = int(nz), int(nx), int(nsweep)
    if case == "a":
        # testing/test2d.F90:40-46 velocity model
        z1 = (nz - 1) * dz
        vel = np.zeros((nz - 1) * (nx - 1))
        for ix in range(1, nx):
            for iz in range(1, nz):
                z = (iz - 1) * dz
                vel[(ix - 1) * (nz - 1) + iz - 1] = \
                    3000.0 + (5000.0 - 3000.0) / z1 * z
    else:
        vel = fixtures["t2d_b_vel"]
    slow = orc.cell_slowness(vel, nz, nx)
    src = orc.source_locate((nz, nx, 1, dz, dx, 1.0, 0.0, 0.0, 0.0), zs, xs,
                            slow=slow, is3d=False)
    tt = orc.solve2d(slow, nz, nx, dz, dx, src, nsweep, eps)
    np.testing.assert_array_equal(tt, fixtures["t2d_%s_tt" % case])
    # the reference's FSM driver (solver2D.F90:1042-1280) reaches the same
    # field: it updates the same masked nodes in a lexicographic order
    np.testing.assert_array_equal(tt, fixtures["t2d_%s_fsm_tt" % case])
    zr, xr = fixtures["t2d_%s_rec" % case]
    ttr = orc.receivers(tt, (nz, nx, 1, dz, dx, 1.0, 0.0, 0.0, 0.0), zr, xr,
                        is3d=False)
    np.testing.assert_array_equal(ttr, fixtures["t2d_%s_ttr" % case])
    if case == "a":
        # the reference's own known-answer test: LSM vs original <= 1e-10
        assert fixtures["t2d_a_orig_maxdiff"][0] <= 1e-10
        np.testing.assert_allclose(tt[::97], fixtures["t2d_a_orig_sample"],
                                   rtol=0, atol=1e-10)
    else:
        # solver2D.F90:1363-1374 (XZ cell averages)
        t = tt.reshape(nx, nz)
        cell = 0.25 * (((t[:-1, :-1] + t[:-1, 1:]) + t[1:, :-1]) + t[1:, 1:])
        np.testing.assert_array_equal(cell.T.ravel(),
                                      fixtures["t2d_b_ttcell_xz"])


# -------------------------------------------------------------- analytic
def test_analytic(fixtures):
    nz, nx, ny, dz, dx, dy, zs, xs, ys, v0, v1 = fixtures["ana_cfg"]
    nz, nx, ny = int(nz), int(nx), int(ny)
    a = orc.analytic_gradient(nz, nx, ny, dz, dx, dy, 0.0, 0.0, 0.0, zs, xs,
                              ys, v0, v1)
    nat = orc.xfast_to_natural(a, nz, nx, ny)
    np.testing.assert_allclose(nat, fixtures["ana_grad_natural"], rtol=1e-15,
                               atol=0)
    np.testing.assert_allclose(a, fixtures["ana_grad_xyz"], rtol=1e-15,
                               atol=0)
    c = orc.analytic_constant(nz, nx, ny, dz, dx, dy, 0.0, 0.0, 0.0, zs, xs,
                              ys, 4200.0)
    np.testing.assert_array_equal(orc.xfast_to_natural(c, nz, nx, ny),
                                  fixtures["ana_const_natural"])


# --------------------------------------------------------------- locator
@pytest.mark.parametrize("tag", ["small", "mid"])
def test_locator_matches_reference(fixtures, tag):
    nz, nx, ny, h, nsta, nev = (int(v) for v in fixtures["loc_%s_cfg" % tag])
    ngrd = nz * nx * ny
    loc = orc.Locator(nev, nsta, ngrd)
    for s, t in enumerate(fixtures["loc_%s_tables" % tag]):
        loc.set_table(s + 1, t)
    for e in range(nev):
        loc.set_observation(e + 1, np.arange(1, nsta + 1),
                            fixtures["loc_%s_picks" % tag][e],
                            fixtures["loc_%s_wts" % tag][e])
    res = fixtures["loc_%s_result" % tag]
    for e in range(nev):
        opt, t0, ob = loc.locate(e + 1, accumulate_t0=True, pad_argmin=True)
        assert opt == int(res[e, 0]), (e, opt, res[e])
        assert t0 == res[e, 1] and ob == res[e, 2]


def test_locator_fresh_state_first_event_and_truth(fixtures):
    """The first event does not depend on the t0 carry-over; with fresh
    state every noise-free event returns its true cell (loctest.c)."""
    nz, nx, ny, h, nsta, nev = (int(v) for v in fixtures["loc_mid_cfg"])
    ngrd = nz * nx * ny
    loc = orc.Locator(nev, nsta, ngrd)
    for s, t in enumerate(fixtures["loc_mid_tables"]):
        loc.set_table(s + 1, t)
    truth = fixtures["loc_mid_true"]
    for e in range(nev):
        loc.set_observation(e + 1, np.arange(1, nsta + 1),
                            fixtures["loc_mid_picks"][e],
                            fixtures["loc_mid_wts"][e])
        opt, t0, ob = loc.locate(e + 1, accumulate_t0=False, pad_argmin=False)
        if e == 0:
            assert opt == int(fixtures["loc_mid_result"][0, 0])
        if e % 2 == 0:          # noise-free events
            assert opt == truth[e] + 1


# ------------------------------------------------ 3D levels and masks
@pytest.mark.parametrize("sweep", range(1, 9))
def test_levels_brute_force(sweep):
    """graph3d.F90:122-353: every node once; level = longest upwind path;
    each level sorted by linear index; nLevels = nx + ny + nz - 2."""
    nz, nx, ny = 5, 4, 6
    ptr, ijkv = orc.levels3d(sweep, nz, nx, ny)
    nlev = nx + ny + nz - 2
    assert ptr[0] == 1 and ptr[-1] == nz * nx * ny + 1 and len(ptr) == nlev + 1
    q = ijkv.reshape(-1, 4)
    assert sorted(q[:, 3]) == list(range(1, nz * nx * ny + 1))
    sz, sx, sy = [(1, 1, 1), (1, -1, 1), (1, 1, -1), (1, -1, -1),
                  (-1, 1, 1), (-1, -1, 1), (-1, 1, -1), (-1, -1, -1)][sweep - 1]
    # longest path from the sweep's first corner along upwind steps
    depth = {}
    for lev in range(nlev):
        nodes = q[ptr[lev] - 1:ptr[lev + 1] - 1]
        assert list(nodes[:, 3]) == sorted(nodes[:, 3])
        for i, j, k, g in nodes:
            best = 0
            for di, dj, dk in [(sz, 0, 0), (0, sx, 0), (0, 0, sy)]:
                u = (i - di, j - dj, k - dk)
                if u in depth:
                    best = max(best, depth[u] + 1)
            depth[(i, j, k)] = best
            assert best == lev


def test_init_masks_box():
    """solver3D.F90:493-565 closed form for a source at (4, 5, 3)."""
    nz, nx, ny = 9, 8, 7
    lim = orc.sweep_limits(1, 1, nz, nx, ny, 4, 5, 3)
    assert list(lim) == [4, 9, 5, 8, 3, 7]
    lim = orc.sweep_limits(8, 1, nz, nx, ny, 4, 5, 3)
    assert list(lim) == [5, 1, 6, 1, 4, 1]
    lim = orc.sweep_limits(2, 0, nz, nx, ny, 4, 5, 3)
    assert list(lim) == [2, 9, 7, 1, 2, 7]


# ------------------------------------------------------------- 3D solve
def _hetero(nz, nx, ny, h, seed):
    from tests.golden.gen_golden_models import hetero_model3d
    return hetero_model3d(np.random.default_rng(seed), nz, nx, ny, h)


def test_level_order_invariance():
    """Level-synchronous (OpenMP) and lexicographic schedules give the same
    field bit for bit: nodes on a level never read each other."""
    nz, nx, ny, h = 23, 19, 17, 15.0
    slow = 1.0 / _hetero(nz, nx, ny, h, 3)
    src = orc.source_locate(grid9(nz, nx, ny, h), 7.7 * h, 10.2 * h, 5.5 * h,
                            slow)
    a = orc.solve3d(slow, nz, nx, ny, h, h, h, src, 2, 4.0, schedule=0)
    b = orc.solve3d(slow, nz, nx, ny, h, h, h, src, 2, 4.0, schedule=1)
    np.testing.assert_array_equal(a, b)
    assert a.max() < HUGE


def test_homogeneous_near_source_exact():
    """Point source in a homogeneous medium: inside the spherical radius the
    factored operators reproduce szero*r up to rounding."""
    n, h, v = 21, 10.0, 3000.0
    slow = np.full((n - 1) ** 3, 1.0 / v)
    src = orc.source_locate(grid9(n, n, n, h), 10.0 * h, 10.0 * h, 10.0 * h,
                            slow)
    tt = orc.solve3d(slow, n, n, n, h, h, h, src, 1, 5.0)
    iz, ix, iy = np.meshgrid(np.arange(n), np.arange(n), np.arange(n),
                             indexing="ij")
    zsa, xsa, ysa = src.zsa - 1, src.xsa - 1, src.ysa - 1
    r = np.sqrt(((iz - zsa) * h) ** 2 + ((ix - xsa) * h) ** 2
                + ((iy - ysa) * h) ** 2)
    exact = (r / v).transpose(2, 1, 0).ravel()     # natural order
    near = (r.transpose(2, 1, 0).ravel() <= 2.0 * h)
    np.testing.assert_allclose(tt[near], exact[near], rtol=1e-12, atol=1e-15)
    # far from the source the plane-wave operators stay first-order accurate
    assert np.abs(tt - exact).max() / exact.max() < 0.02


def test_linear_gradient_vs_analytic():
    """Config-1 physics on a smaller grid: v(z)=3000+0.5 z against the
    closed form of analytic.F90:456-528."""
    n, h = 41, 25.0
    zc = (np.arange(n - 1) + 0.5) * h
    vel = np.tile(3000.0 + 0.5 * zc, (n - 1) * (n - 1))
    slow = orc.cell_slowness(vel, n, n, n)
    zs = xs = ys = 20 * h
    src = orc.source_locate(grid9(n, n, n, h), zs, xs, ys, slow)
    tt = orc.solve3d(slow, n, n, n, h, h, h, src, 2, 5.0)
    vbot = 3000.0 + 0.5 * (n - 1) * h
    a = orc.xfast_to_natural(
        orc.analytic_gradient(n, n, n, h, h, h, 0.0, 0.0, 0.0, zs, xs, ys,
                              3000.0, vbot), n, n, n)
    rel = np.abs(tt - a) / np.maximum(a, 1e-12)
    assert np.median(rel) < 2e-3 and rel.max() < 1e-2


@pytest.mark.skipif(not orc.have_ref(), reason="reference build absent")
def test_against_original_fteik3d():
    """refsrc/fteik.f90 (Noble & Gesret) solves the same discretisation with
    its own gathers and its own operator code (and rounds its seeds to
    single precision); the level-set solver agrees with it to ~1e-7, far
    below what a wrong gather, mask or sweep order would give."""
    import ctypes
    n, h = 25, 20.0
    nz, nx, ny = n, n - 2, n - 4
    vel = _hetero(nz, nx, ny, h, 5)
    slow = 1.0 / vel
    zs, xs, ys = 9.3 * h, 8.6 * h, 7.7 * h
    src = orc.source_locate(grid9(nz, nx, ny, h), zs, xs, ys, slow)
    tt = orc.solve3d(slow, nz, nx, ny, h, h, h, src, 2, 3.0)
    orig = np.zeros(nz * nx * ny)
    R = orc.ref()
    R.refh_orig_solver3d(orc._d(np.ascontiguousarray(vel)), orc._d(orig),
                         ctypes.c_int(nz), ctypes.c_int(nx), ctypes.c_int(ny),
                         ctypes.c_double(zs), ctypes.c_double(xs),
                         ctypes.c_double(ys), ctypes.c_double(h),
                         ctypes.c_double(h), ctypes.c_double(h),
                         ctypes.c_int(2), ctypes.c_double(3.0))
    rel = np.abs(tt - orig) / np.maximum(orig, 1e-12)
    assert np.sqrt(np.mean(rel ** 2)) < 1e-6
    assert rel.max() < 1e-5
