"""FP64 CPU references of the equation-of-state surrogate (test helper).

`forward` / `backward` are the exact network and its hand-written gradient;
with `bf16=True` they round every product operand to BF16 (round to nearest
even of the FP32 value) at the points where the device kernels do, so the
device results can be compared operand for operand.  `bound` is the
rigorous first-order BF16 error bound of the device forward pass against the
exact network.
"""
import numpy as np

from implicit_dpd_amd.surrogate import unpack

U_BF16 = 2.0 ** -8  # unit roundoff of BF16 (8 significand bits)
SILU_LIP = 1.0998   # max |d silu / dz|


def bf16(a):
    """Round to BF16 (nearest even) through FP32, returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32)
    return (r << 16).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, x, bf=False):
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = (np.asarray(a, np.float64) for a in unpack(p))
    x = np.asarray(x, np.float64)
    z1 = r(x) @ r(W1).T + b1
    h1 = silu(z1)
    z2 = r(h1) @ r(W2).T + b2
    h2 = silu(z2)
    y = r(h2) @ r(W3).T + b3
    return y, (x, z1, h1, z2, h2)


def loss(p, x, t, bf=False):
    y, _ = forward(p, x, bf)
    return float(((y - np.asarray(t, np.float64)) ** 2).sum() / (2.0 * len(x)))


def backward(p, x, t, bf=False):
    """Gradient of loss = sum (y - t)^2 / (2n) w.r.t. the flat parameters."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = (np.asarray(a, np.float64) for a in unpack(p))
    n = len(x)
    y, (x, z1, h1, z2, h2) = forward(p, x, bf)
    dy = (y - np.asarray(t, np.float64)) / n
    if bf:
        dy = np.asarray(np.float32(1.0) * dy.astype(np.float32), np.float64)
    dz2 = (r(dy) @ r(W3)) * dsilu(z2)
    dz1 = (r(dz2) @ r(W2)) * dsilu(z1)
    g = np.zeros(sum(a.size for a in (W1, b1, W2, b2, W3, b3)))
    parts = [r(dz1).T @ r(x), r(dz1).sum(0), r(dz2).T @ r(h1), r(dz2).sum(0),
             r(dy).T @ r(h2), r(dy).sum(0)]
    off = 0
    for a in parts:
        g[off:off + a.size] = a.ravel()
        off += a.size
    return g


def bound(p, x):
    """|y_device - y_exact| <= bound, elementwise (first order in U_BF16,
    with the FP32 accumulation error folded into a 1e-6 relative term)."""
    W1, b1, W2, b2, W3, b3 = (np.abs(np.asarray(a, np.float64)) for a in unpack(p))
    _, (x, z1, h1, z2, h2) = forward(p, x, False)
    e = 2 * U_BF16 + U_BF16 ** 2
    ax, ah1, ah2 = np.abs(x), np.abs(h1), np.abs(h2)
    e1 = e * (ax @ W1.T) + 1e-6 * (ax @ W1.T + b1)
    eh1 = SILU_LIP * e1
    e2 = (1 + U_BF16) * (eh1 @ W2.T) + e * (ah1 @ W2.T) + 1e-6 * (ah1 @ W2.T + b2)
    eh2 = SILU_LIP * e2
    ey = (1 + U_BF16) * (eh2 @ W3.T) + e * (ah2 @ W3.T) + 1e-6 * (ah2 @ W3.T + b3)
    return ey
