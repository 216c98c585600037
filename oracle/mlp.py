"""CPU check for the locator surrogate's training kernels (test
infrastructure only; nothing in the product imports it).

The surrogate is not part of the reference (locate.F90 searches the full
grid), so this is not a restatement of reference code: it is a float64
numpy statement of the network fteik_amd trains,
3 -> 256 -> 256 -> 256 -> 256 -> ntf with SiLU, with the GPU's BF16
rounding applied at the same points (activations and weights entering the
MFMA GEMMs, the output gradient, hidden-layer gradients), so the device
forward pass and parameter gradients can be compared with it to the
accuracy of FP32 accumulation.

Parameter vector layout (as fteik_surrogate_getParameters returns it):
W1 [256 x 3], b1 [256], then for the three 256 x 256 layers W [256 x 256]
(row-major [out][in]) and b [256], then W5 [ntf x 256], b5 [ntf].
"""
import numpy as np

H = 256


def bf16(a):
    """Round to bfloat16 (nearest even) and back, as the device does."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p, ntf):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    out = {}

    def take(name, shape):
        nonlocal o
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    take("W1", (H, 3))
    take("b1", (H,))
    for l in range(3):
        take("W%d" % (l + 2), (H, H))
        take("b%d" % (l + 2), (H,))
    take("W5", (ntf, H))
    take("b5", (ntf,))
    assert o == p.size
    return out


def pack(d, ntf):
    parts = [d["W1"], d["b1"]]
    for l in range(3):
        parts += [d["W%d" % (l + 2)], d["b%d" % (l + 2)]]
    parts += [d["W5"], d["b5"]]
    return np.concatenate([np.ravel(x) for x in parts])


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, x, ntf):
    """Returns the outputs and the pre-activations of the 4 hidden layers."""
    P = unpack(p, ntf)
    x = np.asarray(x, dtype=np.float32).astype(np.float64)
    z = [x @ P["W1"].T.astype(np.float32).astype(np.float64) + P["b1"]]
    h = [bf16(_silu(z[0]))]
    for l in range(3):
        W = bf16(P["W%d" % (l + 2)])
        z.append(h[-1] @ W.T + P["b%d" % (l + 2)])
        h.append(bf16(_silu(z[-1])))
    y = h[-1] @ bf16(P["W5"]).T + P["b5"]
    return y, z, h


def gradients(p, x, tgt, ntf):
    """Mean-squared-error loss and parameter gradients, with the device's
    BF16 rounding of dy and of the hidden-layer gradients."""
    P = unpack(p, ntf)
    x = np.asarray(x, dtype=np.float32).astype(np.float64)
    y, z, h = forward(p, x, ntf)
    n = y.size
    d = y - tgt
    loss = np.sum(d * d) / n
    dy = bf16(2.0 * d / n)
    g = {}
    g["W5"] = dy.T @ h[3]
    g["b5"] = dy.sum(axis=0)
    dh = dy @ bf16(P["W5"])
    dz = bf16(dh * _dsilu(z[3]))
    for l in (2, 1, 0):
        g["W%d" % (l + 2)] = dz.T @ h[l]
        g["b%d" % (l + 2)] = dz.sum(axis=0)
        dh = dz @ bf16(P["W%d" % (l + 2)])
        dz = bf16(dh * _dsilu(z[l]))
    g["W1"] = dz.T @ x
    g["b1"] = dz.sum(axis=0)
    return loss, pack(g, ntf)
