"""The CPU statement of the locator surrogate (oracle/mlp.py) is checked
by finite differences: without the BF16 roundings its analytic parameter
gradient equals central differences of its loss."""
import numpy as np

from oracle import mlp


def test_cpu_statement_gradients_by_finite_differences():
    ntf, n = 3, 9
    rng = np.random.default_rng(2)
    npar = 3 * 256 + 256 + 3 * (256 * 256 + 256) + ntf * 256 + ntf
    p = 0.05 * rng.standard_normal(npar)
    x = rng.uniform(-1, 1, (n, 3))
    tgt = rng.standard_normal((n, ntf))
    saved = mlp.bf16
    try:
        mlp.bf16 = lambda a: np.asarray(a, dtype=np.float64)
        _, g = mlp.gradients(p, x, tgt, ntf)
        idx = rng.choice(npar, 40, replace=False)
        for i in idx:
            e = np.zeros(npar)
            e[i] = 1e-6
            lp, _ = mlp.gradients(p + e, x, tgt, ntf)
            lm, _ = mlp.gradients(p - e, x, tgt, ntf)
            fd = (lp - lm) / 2e-6
            assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(g[i]), (i, fd, g[i])
    finally:
        mlp.bf16 = saved


