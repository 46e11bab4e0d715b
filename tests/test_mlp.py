"""Equation-of-state surrogate: CPU pins of the FP64 reference, then the
device network against it.

Tolerances: the reference gradient equals FP64 autograd to 1e-10 and central
differences to 1e-6; the device forward pass stays inside the rigorous BF16
error bound of the exact network, and matches the BF16-emulating reference
to 1e-4 of the output scale (rare BF16 rounding flips of FP32-vs-FP64
pre-activations allowed); device gradients match the BF16-emulating
reference to 1e-2 and exact autograd to 5e-2 (relative Frobenius norm per
tensor); training is bit-reproducible.
"""
import numpy as np
import pytest

from implicit_dpd_amd import surrogate as S

import mlp_ref as R

TENSORS = [("W1", S.OFF_W1, S.OFF_B1), ("b1", S.OFF_B1, S.OFF_W2), ("W2", S.OFF_W2, S.OFF_B2),
           ("b2", S.OFF_B2, S.OFF_W3), ("W3", S.OFF_W3, S.OFF_B3), ("b3", S.OFF_B3, S.N_PARAMS)]


def _params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(S.N_PARAMS)
    for name, a, b in TENSORS:
        fan_in = 4 if name in ("W1", "b1") else 128
        p[a:b] = rng.uniform(-1, 1, b - a) / np.sqrt(fan_in)
    return p.astype(np.float32)


def _data(seed, n):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.5, 1.5, (n, 4))
    t = np.stack([np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2], np.cos(x[:, 3]) - x[:, 0] ** 2], 1)
    return x.astype(np.float32), t.astype(np.float32)


def _torch_grad(p, x, t):
    import torch
    pt = torch.tensor(np.asarray(p, np.float64), requires_grad=True)
    W = [pt[a:b] for _, a, b in TENSORS]
    W1, b1 = W[0].reshape(128, 4), W[1]
    W2, b2 = W[2].reshape(128, 128), W[3]
    W3, b3 = W[4].reshape(2, 128), W[5]
    xt = torch.tensor(np.asarray(x, np.float64))
    tt = torch.tensor(np.asarray(t, np.float64))
    h1 = torch.nn.functional.silu(xt @ W1.T + b1)
    h2 = torch.nn.functional.silu(h1 @ W2.T + b2)
    y = h2 @ W3.T + b3
    L = ((y - tt) ** 2).sum() / (2 * len(x))
    L.backward()
    return L.item(), pt.grad.numpy()


# ---------------------------------------------------------------- CPU pins
def test_bf16_rounding_known_answers():
    one = 1.0
    # BF16 spacing at 1 is 2^-7, so 1 + 2^-8 is a tie
    assert R.bf16(one + 2.0 ** -8) == 1.0                     # tie -> even (down)
    assert R.bf16(one + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6     # tie -> even (up)
    assert R.bf16(one + 2.0 ** -8 + 2.0 ** -20) == 1.0 + 2.0 ** -7
    assert R.bf16(one + 2.0 ** -9) == 1.0
    assert R.bf16(-2.5) == -2.5 and R.bf16(0.0) == 0.0
    assert abs(R.bf16(np.pi) - np.pi) <= R.U_BF16 * np.pi


def test_reference_gradient_matches_autograd_and_differences():
    p = _params(1).astype(np.float64)
    x, t = _data(2, 64)
    g = R.backward(p, x, t)
    L, gt = _torch_grad(p, x, t)
    assert abs(R.loss(p, x, t) - L) <= 1e-12 * abs(L)
    assert np.abs(g - gt).max() <= 1e-10 * np.abs(gt).max()
    rng = np.random.default_rng(3)
    for _ in range(4):
        v = rng.normal(size=p.size)
        v /= np.linalg.norm(v)
        h = 1e-5
        fd = (R.loss(p + h * v, x, t) - R.loss(p - h * v, x, t)) / (2 * h)
        assert abs(fd - g @ v) <= 1e-6 * max(abs(fd), 1e-3), (fd, g @ v)


def test_bound_covers_emulated_rounding():
    p = _params(4)
    x, _ = _data(5, 256)
    y0, _ = R.forward(p, x, False)
    yb, _ = R.forward(p, x, True)
    assert np.all(np.abs(yb - y0) <= R.bound(p, x))


# ---------------------------------------------------------------- device
@pytest.mark.gpu
def test_device_forward_within_bf16_bound():
    net = S.Surrogate(max_batch=512, seed=7)
    p = net.params
    x, _ = _data(6, 1000)  # two device batches
    y = net.forward(x)
    y0, _ = R.forward(p, x, False)
    yb, _ = R.forward(p, x, True)
    assert np.all(np.abs(y - y0) <= R.bound(p, x))
    scale = np.abs(y0).max()
    assert np.abs(y - yb).max() <= 1e-4 * scale, np.abs(y - yb).max()
    net.close()


@pytest.mark.gpu
def test_device_params_roundtrip_and_forward():
    net = S.Surrogate(max_batch=256, seed=1)
    p = _params(8)
    net.params = p
    assert np.array_equal(net.params, p)
    x, _ = _data(9, 100)
    yb, _ = R.forward(p, x, True)
    assert np.abs(net.forward(x) - yb).max() <= 1e-4 * np.abs(yb).max()
    net.close()


@pytest.mark.gpu
def test_device_gradients():
    net = S.Surrogate(max_batch=1024, seed=3)
    p = _params(10)
    net.params = p
    x, t = _data(11, 1000)
    L, g = net.loss_grad(x, t)
    gb = R.backward(p, x, t, bf=True)
    Lx, gx = _torch_grad(p, x, t)
    assert abs(L - R.loss(p, x, t, bf=True)) <= 1e-4 * Lx
    assert abs(L - Lx) <= 2e-2 * Lx
    for name, a, b in TENSORS:
        nb = np.linalg.norm(gb[a:b])
        assert np.linalg.norm(g[a:b] - gb[a:b]) <= 1e-2 * nb + 1e-9, name
        assert np.linalg.norm(g[a:b] - gx[a:b]) <= 5e-2 * np.linalg.norm(gx[a:b]) + 1e-9, name
    net.close()


@pytest.mark.gpu
def test_device_training_converges_and_is_reproducible():
    x, t = _data(12, 4096)
    out = []
    for _ in range(2):
        net = S.Surrogate(max_batch=4096, seed=5)
        L0, _ = net.loss_grad(x, t)
        hist = net.train(x, t, epochs=40, batch=256, lr=3e-3, seed=9)
        L1, _ = net.loss_grad(x, t)
        out.append((net.params, hist, L0, L1))
        net.close()
    (p1, h1, L0, L1), (p2, h2, _, _) = out
    assert np.array_equal(p1, p2) and np.array_equal(h1, h2)
    assert L1 < 0.02 * L0, (L0, L1)
    assert h1[-1] < h1[0]


# ---------------------------------------------------------------- calibration
def test_slab_density_estimator():
    from implicit_dpd_amd.calibrate import slab_rho
    prof = np.zeros(40)
    prof[30:40] = 6.0
    prof[0:6] = 6.0          # slab wrapping through z = 0
    prof[29] = prof[6] = 3.0  # interface bins
    assert abs(slab_rho(prof, 1.0) - 6.0) < 1e-12
    prof2 = prof * 2.0
    assert abs(slab_rho(prof2, 2.0) - 6.0) < 1e-12
    # no dilute vapour (the liquid filled the box or never condensed): the
    # homogeneous fluid's density, the mean over the box
    flat = np.full(40, 3.0)
    flat[10:20] = 4.5
    assert abs(slab_rho(flat, 1.0) - flat.mean()) < 1e-12
    # the slab estimate holds while the vapour stays below a quarter of the
    # liquid's density, and gives way to the box mean above that
    lo = [slab_rho(np.where(np.arange(40) < 20, 6.0, v), 1.0) for v in (0.5, 1.0, 1.4)]
    assert lo[0] == lo[1] == lo[2] == 6.0
    assert slab_rho(np.where(np.arange(40) < 20, 6.0, 1.6), 1.0) == 3.8


@pytest.mark.gpu
def test_calibration_pipeline_small():
    from implicit_dpd_amd import calibrate as K
    X, Y = K.ensemble_samples(1, n_rep=4, nmin=2000, nmax=3000, steps=200, every=50)
    assert X.shape == (4, 4) and Y.shape == (4, 2) and np.all(np.isfinite(Y))
    rho, pzz = K.plain_run(-40.0, n=20000, steps=400)
    assert 3.0 < rho < 9.0 and np.isfinite(pzz)
