"""GPU checks of the locator surrogate: the BF16 MFMA GEMM against numpy,
the network's forward pass and parameter gradients against the float64
CPU statement (oracle/mlp.py), and training plus surrogate-pruned location
against the full grid search on a small problem."""
import ctypes
from ctypes import byref, c_double, c_int

import numpy as np
import pytest

from oracle import mlp
from tests.gpu_helpers import lib, _d, _i

pytestmark = pytest.mark.gpu

_fp = ctypes.POINTER(ctypes.c_float)


def _f(a):
    return a.ctypes.data_as(_fp)


@pytest.mark.parametrize("M,N,K", [(64, 64, 32), (100, 70, 50),
                                   (257, 48, 256), (33, 256, 3),
                                   (1000, 256, 256)])
def test_gemm_bf16_matches_numpy(M, N, K):
    rng = np.random.default_rng(M * 7 + N)
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    C = np.zeros((M, N), dtype=np.float32)
    ierr = c_int(1)
    lib().fteik_amd_gemmBf16(M, N, K, _f(A), _f(B), _f(C), byref(ierr))
    assert ierr.value == 0
    ref = mlp.bf16(A) @ mlp.bf16(B)
    # FP32 accumulation of K products
    tol = 4e-7 * K * np.abs(mlp.bf16(A)) @ np.abs(mlp.bf16(B)) + 1e-30
    assert np.all(np.abs(C - ref) <= tol), np.max(np.abs(C - ref) / tol)


@pytest.mark.parametrize("R", [32, 45, 1000])
def test_fused_hidden_layer(R):
    """The fused inversion layer equals the GEMM + activation kernels bit
    for bit, and the float64 statement of the layer on the BF16 inputs to
    BF16 rounding (silu'(z_0) z_d for the tangent blocks)."""
    rng = np.random.default_rng(R)
    H = 256
    hin = rng.standard_normal((4 * R, H)).astype(np.float32)
    W = (rng.standard_normal((H, H)) / 16).astype(np.float32)
    b = (0.1 * rng.standard_normal(H)).astype(np.float32)
    fused = np.zeros((4 * R, H), dtype=np.uint16)
    unfused = np.zeros((4 * R, H), dtype=np.uint16)
    ierr = c_int(1)
    up = ctypes.POINTER(ctypes.c_ushort)
    lib().fteik_amd_invHiddenBf16(R, _f(hin), _f(W), _f(b),
                                  fused.ctypes.data_as(up),
                                  unfused.ctypes.data_as(up), byref(ierr))
    assert ierr.value == 0
    np.testing.assert_array_equal(fused, unfused)
    got = (fused.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    z = mlp.bf16(hin).astype(np.float64) @ mlp.bf16(W).astype(np.float64).T
    z0 = z[:R] + b
    sig = 1.0 / (1.0 + np.exp(-z0))
    want = np.concatenate([z0 * sig] + [sig * (1 + z0 * (1 - sig)) * z[d * R:(d + 1) * R]
                                        for d in (1, 2, 3)])
    tol = 1e-2 * np.abs(want) + 1e-3 * np.abs(want).max()
    assert np.all(np.abs(got - want) <= tol), np.max(np.abs(got - want) / tol)


def _init_surrogate(ntf, seed=3):
    L = lib()
    ierr = c_int(1)
    z = np.zeros(ntf)
    x = np.linspace(100.0, 900.0, ntf)
    y = np.linspace(50.0, 700.0, ntf)
    L.fteik_surrogate_initialize(ntf, 20, 21, 22, 50.0, 50.0, 50.0, _d(z),
                                 _d(x), _d(y), 4000.0, 1.0, seed, byref(ierr))
    assert ierr.value == 0
    np_ = c_int()
    L.fteik_surrogate_getNumberOfParameters(byref(np_), byref(ierr))
    return L, np_.value


def test_parameter_count():
    L, n = _init_surrogate(48)
    assert n == 210736          # 3 -> 256 x 4 -> 48
    L.fteik_surrogate_free()


@pytest.mark.parametrize("ntf,n", [(5, 37), (48, 4100)])
def test_forward_and_gradients_match_cpu_statement(ntf, n):
    """n = 4100 takes the training kernels' multi-part paths: the weight
    gradients split over 15 row ranges, the bias gradients over 65 row
    blocks (the last one of 4 rows), all summed by the Adam kernel."""
    L, npar = _init_surrogate(ntf)
    ierr = c_int(1)
    rng = np.random.default_rng(11)
    # biases non-zero so every term is exercised
    p = np.zeros(npar, dtype=np.float32)
    L.fteik_surrogate_getParameters(npar, _f(p), byref(ierr))
    P = mlp.unpack(p, ntf)
    for k in ("b1", "b2", "b3", "b4", "b5"):
        P[k][:] = 0.1 * rng.standard_normal(P[k].shape)
    p = mlp.pack(P, ntf).astype(np.float32)
    L.fteik_surrogate_setParameters(npar, _f(p), byref(ierr))
    assert ierr.value == 0
    x = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    tgt = rng.standard_normal((n, ntf)).astype(np.float32)
    y = np.zeros((n, ntf), dtype=np.float32)
    L.fteik_surrogate_forward(n, _f(x), _f(y), byref(ierr))
    assert ierr.value == 0
    yr, _, _ = mlp.forward(p, x, ntf)
    scale = np.abs(yr).max()
    assert np.abs(y - yr).max() <= 1e-2 * scale
    g = np.zeros(npar, dtype=np.float32)
    loss = c_double()
    L.fteik_surrogate_gradients(n, _f(x), _f(tgt), _f(g), byref(loss),
                                byref(ierr))
    assert ierr.value == 0
    lr, gr = mlp.gradients(p, x, tgt.astype(np.float64), ntf)
    assert abs(loss.value - lr) <= 1e-3 * abs(lr)
    G, R = mlp.unpack(g, ntf), mlp.unpack(gr, ntf)
    for k in R:
        err = np.linalg.norm(G[k] - R[k]) / (np.linalg.norm(R[k]) + 1e-30)
        assert err < 2e-2, (k, err)
    L.fteik_surrogate_free()


def test_training_and_pruned_location():
    """Tables from the batched solver, a short training run, then the
    surrogate-pruned search agrees with the full grid search."""
    L = lib()
    ierr = c_int(1)
    nz, nx, ny, h = 30, 40, 36, 200.0
    zc = (np.arange(nz - 1) + 0.5) * h
    vel = np.ascontiguousarray(np.tile(4000.0 + 0.4 * zc, (nx - 1) * (ny - 1)))
    rng = np.random.default_rng(21)
    nsta = 12
    sx = rng.uniform(0, (nx - 3) * h, nsta)
    sy = rng.uniform(0, (ny - 3) * h, nsta)
    sz = np.zeros(nsta)
    L.fteik_batch3d_initialize(nsta, byref(ierr))
    for k in range(nsta):
        L.fteik_batch3d_setGrid(k + 1, nz, nx, ny, 0.0, 0.0, 0.0, h, h, h,
                                _d(vel), 0.0, sx[k], sy[k], byref(ierr))
        assert ierr.value == 0
    L.fteik_batch3d_solve(2, 5.0, byref(ierr))
    assert ierr.value == 0
    ngrd = nz * nx * ny
    nev = 300
    L.locate_initialize(nev, nsta, ngrd, byref(ierr))
    L.locate_setReferenceCompatibility(0, byref(ierr))
    tabs = np.zeros((nsta, ngrd))
    for k in range(nsta):
        L.locate_setTravelTimeFieldFromBatch3d(k + 1, k + 1, byref(ierr))
        assert ierr.value == 0
        t = np.zeros(ngrd)
        L.fteik_batch3d_getTravelTimeField64f(k + 1, ngrd, 0, _d(t),
                                              byref(ierr))
        tabs[k] = t
    L.fteik_batch3d_free()
    hyp = rng.integers(0, ngrd, nev)
    picks = tabs[:, hyp].T + rng.uniform(0, 50, nev)[:, None] + \
        rng.normal(0, 0.02, (nev, nsta))
    o2t = np.arange(1, nsta + 1, dtype=np.int32)
    w = np.ones(nsta)
    for e in range(nev):
        p = np.ascontiguousarray(picks[e])
        L.locate_setObservation64f(e + 1, nsta, False, 0.0, _i(o2t), _d(p),
                                   _d(w), byref(ierr))
    L.fteik_surrogate_initialize(nsta, nz, nx, ny, h, h, h, _d(sz), _d(sx),
                                 _d(sy), 4500.0, 1.0, 13, byref(ierr))
    assert ierr.value == 0
    l0, l1 = c_double(), c_double()
    L.fteik_surrogate_train(5, 2048, 1e-3, 200000, byref(l0), byref(ierr))
    assert ierr.value == 0
    L.fteik_surrogate_train(1500, 2048, 2e-3, 200000, byref(l1), byref(ierr))
    assert ierr.value == 0
    print("surrogate loss %.3g -> %.3g" % (l0.value, l1.value))
    assert l1.value < 0.05 * l0.value
    ev = np.arange(1, nev + 1, dtype=np.int32)
    full = [np.zeros(nev, dtype=np.int32), np.zeros(nev), np.zeros(nev)]
    L.locate_locateEvents(nev, _i(ev), _i(full[0]), _d(full[1]), _d(full[2]),
                          byref(ierr))
    assert ierr.value == 0
    sur = [np.zeros(nev, dtype=np.int32), np.zeros(nev), np.zeros(nev)]
    nfb = c_int()
    L.locate_locateEventsSurrogate(nev, _i(ev), 8, 6, _i(sur[0]), _d(sur[1]),
                                   _d(sur[2]), byref(nfb), byref(ierr))
    assert ierr.value == 0
    agree = np.mean(sur[0] == full[0])
    print("agreement %.4f, fallbacks %d" % (agree, nfb.value))
    assert agree >= 0.99
    same = sur[0] == full[0]
    np.testing.assert_array_equal(sur[1][same], full[1][same])
    np.testing.assert_array_equal(sur[2][same], full[2][same])
    L.locate_setReferenceCompatibility(1, byref(ierr))
    L.locate_finalize()
    L.fteik_surrogate_free()
