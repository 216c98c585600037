"""Seeded synthetic models shared by gen_golden.py and the tests that
rebuild a fixture's inputs (data generators only; no reference code)."""
import numpy as np


def smooth_noise(rng, shape, sigma):
    """Separable Gaussian smoothing of N(0,1) noise normalised to max|g|=1."""
    g = rng.standard_normal(shape)
    r = int(3 * sigma)
    x = np.arange(-r, r + 1)
    w = np.exp(-0.5 * (x / sigma) ** 2)
    w /= w.sum()
    for ax in range(len(shape)):
        g = np.apply_along_axis(
            lambda v: np.convolve(np.pad(v, r, mode="reflect"), w,
                                  mode="valid"), ax, g)
    return g / np.abs(g).max()


def hetero_model3d(rng, nz, nx, ny, h):
    """v(z) = 2500 + 0.8 z times a smooth +-8 % perturbation; cells in
    natural (z-fastest) order."""
    zc = (np.arange(nz - 1) + 0.5) * h
    v = 2500.0 + 0.8 * zc
    g = smooth_noise(rng, (ny - 1, nx - 1, nz - 1), 2.0)
    return (v[None, None, :] * (1.0 + 0.08 * g)).ravel()
