"""Pins the network oracle (oracle/net.py) on the CPU: its convolution against
scipy.ndimage.correlate, its pooling and upsampling transposes by the adjoint
identity, and its whole backward pass against central finite differences of
its forward pass (so a dropped term, a wrong tap sign or a transposed weight
in the backward restatement fails)."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import net as on


def test_conv_matches_scipy_correlate():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((5, 6, 7, 3))
    W = rng.standard_normal((27, 2, 3))
    b = rng.standard_normal(2)
    Y = on.conv(X, W, b)
    for co in range(2):
        ref = np.full(X.shape[:3], b[co])
        for ci in range(3):
            k = np.zeros((3, 3, 3))   # [dz+1, dy+1, dx+1]
            for t, (dx, dy, dz) in enumerate(on.taps()):
                k[dz + 1, dy + 1, dx + 1] = W[t, co, ci]
            ref += ndimage.correlate(X[..., ci], k, mode="constant", cval=0.0)
        assert np.abs(Y[..., co] - ref).max() < 1e-12


@pytest.mark.parametrize("shape", [(4, 6, 8), (5, 3, 7), (1, 2, 3)])
def test_pool_upsample_adjoints(shape):
    rng = np.random.default_rng(2)
    half = tuple((n + 1) // 2 for n in shape)
    x = rng.standard_normal(shape + (2,))
    y = rng.standard_normal(half + (2,))
    assert abs((on.pool(x) * y).sum() - (x * on.pool_t(y, shape)).sum()) < 1e-12
    assert abs((on.upsample(y, shape) * x).sum() - (y * on.upsample_t(x, half)).sum()) < 1e-12
    # pooling a constant field gives the constant (windows clipped at the faces)
    assert np.allclose(on.pool(np.ones(shape + (1,))), 1.0)


def test_conv_input_and_weight_grads_are_adjoint():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((3, 4, 5, 2))
    W = rng.standard_normal((27, 3, 2))
    G = rng.standard_normal((3, 4, 5, 3))
    lhs = (on.conv(X, W) * G).sum()
    assert abs(lhs - (X * on.conv_input_grad(G, W)).sum()) < 1e-10
    assert abs(lhs - (W * on.conv_weight_grad(X, G)).sum()) < 1e-10


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(4)
    theta = on.random_params(5)
    shape = (3, 4, 5)
    x = rng.standard_normal(shape + (4,))
    t = rng.standard_normal(shape + (3,))
    loss, g = on.loss_and_grad(theta, x, t)
    # every layer's weights and biases, at real (non-padding) entries
    idx, off = [], 0
    for l, (ci, co) in enumerate(zip(on.CIN, on.COUT)):
        for _ in range(4):
            tap, o, i = rng.integers(27), rng.integers(on.COUT_REAL[l]), rng.integers(on.CIN_REAL[l])
            idx.append(off + (tap * co + o) * ci + i)
        off += 27 * ci * co
        idx.append(off + rng.integers(on.COUT_REAL[l]))
        off += co
    h = 1e-6
    for k in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (on.loss_and_grad(tp, x, t)[0] - on.loss_and_grad(tm, x, t)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
    assert np.isfinite(loss)
    # padding entries carry no gradient
    p = on.unpack(g)
    assert np.all(p["E0A"][0][:, :, 4:] == 0) and np.all(p["HEAD"][0][:, 3:, :] == 0)


def test_bf16_rounding():
    a = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1e-30, 3.0e38])
    r = on.bf16(a)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 4 * 2 ** -8 and r[3] == -2.5
    assert abs(r[4] - 1e-30) / 1e-30 < 2 ** -8


def test_adam_oracle_closed_forms():
    """Adam's bias correction makes the first step, and every step of a
    constant gradient, exactly lr g / (|g| + eps): m_hat = g and v_hat = g^2.
    A zero gradient on the first step leaves the weights; the update scales
    with lr and, for |g| >> eps, does not depend on the gradient's scale."""
    rng = np.random.default_rng(3)
    n = 1000
    w0 = rng.standard_normal(n).astype(np.float32)
    g = (rng.standard_normal(n) * 10.0 ** rng.uniform(-4, 3, n)).astype(np.float32)
    lr = 1e-3
    w, m, v = on.adam_step(w0, np.zeros(n), np.zeros(n), g, 1, lr)
    step = lr * g.astype(np.float64) / (np.abs(g) + 1e-8)
    assert np.allclose(w0 - w, step, rtol=1e-3, atol=1e-9)
    for t in range(2, 6):
        wp = w
        w, m, v = on.adam_step(w, m, v, g, t, lr)
        assert np.allclose(wp - w, step, rtol=1e-3, atol=1e-9)
    w1, _, _ = on.adam_step(w0, np.zeros(n), np.zeros(n), np.zeros(n), 1, lr)
    assert np.array_equal(w1, w0)
    w2, _, _ = on.adam_step(w0, np.zeros(n), np.zeros(n), 1e3 * g, 1, 2 * lr)
    assert np.allclose(w0 - w2, 2 * lr * np.sign(g), rtol=1e-3, atol=1e-9)
    # a sign change of the gradient after a step reverses the momentum only partly
    w3, m3, v3 = on.adam_step(np.zeros(n), np.zeros(n), np.zeros(n), np.ones(n), 1, lr)
    w4, _, _ = on.adam_step(w3, m3, v3, -np.ones(n), 2, lr)
    mh = (0.9 * 0.1 - 0.1) / (1 - 0.81)
    vh = (0.999 * 0.001 + 0.001) / (1 - 0.999 ** 2)
    assert np.allclose(w3 - w4, lr * mh / (np.sqrt(vh) + 1e-8), rtol=1e-4)
