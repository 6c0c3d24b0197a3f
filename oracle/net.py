"""ORACLE — test infrastructure only.

NumPy (FP64) restatement of the learned warm start's convolutional network
(librobtop_amd's rt_net_*; it has no counterpart in the reference). Only
``tests/`` imports this module. With ``emulate=True`` every tensor the GPU
stores as bfloat16 (inputs, weights, activations, backward gradients) is
rounded to bfloat16 at the same point, so the GPU result differs from it only
by FP32 accumulation order.

Tensors are (nz, ny, nx, C), x fastest. A 3x3x3 convolution is the
zero-padded cross-correlation Y[p][co] = b[co] + sum_t sum_ci W[t][co][ci]
X[p + d_t][ci], tap t = (dx+1) + 3 (dy+1) + 9 (dz+1). Network: two
convolutions + ReLU at each of three resolutions (16 / 32 / 64 channels),
2x2x2 average pooling down (windows clipped at the upper faces), nearest
upsampling up, a convolution + additive skip + ReLU and a second convolution
+ ReLU at each decoder level, and a 3x3x3 output convolution. Loss = mean over
positions and the 3 output channels of the squared error.

Pinned by: the convolution against scipy.ndimage.correlate, the pooling and
upsampling transposes by the adjoint identity, and the whole backward pass by
central finite differences (tests/test_net_oracle.py).
"""
import numpy as np

LAYERS = ("E0A", "E0B", "E1A", "E1B", "E2A", "E2B", "D1A", "D1B", "D0A", "D0B", "HEAD")
CIN = (16, 16, 16, 32, 32, 64, 64, 32, 32, 16, 16)
COUT = (16, 16, 32, 32, 64, 64, 32, 32, 16, 16, 16)
CIN_REAL = (4, 16, 16, 32, 32, 64, 64, 32, 32, 16, 16)
COUT_REAL = (16, 16, 32, 32, 64, 64, 32, 32, 16, 16, 3)


def taps():
    return [(t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1) for t in range(27)]


def bf16(a):
    """Round to bfloat16 (nearest even, through binary32), returned as float64."""
    a32 = np.ascontiguousarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(theta):
    """Parameter vector -> {name: (W[27][cout][cin], b[cout])}."""
    out, off = {}, 0
    for name, ci, co in zip(LAYERS, CIN, COUT):
        w = np.asarray(theta[off:off + 27 * ci * co], np.float64).reshape(27, co, ci)
        off += 27 * ci * co
        b = np.asarray(theta[off:off + co], np.float64)
        off += co
        out[name] = (w, b)
    return out


def num_params():
    return sum(27 * ci * co + co for ci, co in zip(CIN, COUT))


def random_params(seed):
    """He-normal weights with non-zero biases (for gradient checks), padding zero."""
    rng = np.random.default_rng(seed)
    theta = np.zeros(num_params())
    off = 0
    for l, (ci, co) in enumerate(zip(CIN, COUT)):
        w = np.zeros((27, co, ci))
        w[:, :COUT_REAL[l], :CIN_REAL[l]] = rng.standard_normal((27, COUT_REAL[l], CIN_REAL[l])) * np.sqrt(
            2.0 / (27 * CIN_REAL[l]))
        theta[off:off + w.size] = w.ravel()
        off += w.size
        b = np.zeros(co)
        b[:COUT_REAL[l]] = 0.05 * rng.standard_normal(COUT_REAL[l])
        theta[off:off + co] = b
        off += co
    return theta


def conv(X, W, b=None):
    nz, ny, nx, _ = X.shape
    Xp = np.pad(X, ((1, 1), (1, 1), (1, 1), (0, 0)))
    Y = np.zeros((nz, ny, nx, W.shape[1]))
    for t, (dx, dy, dz) in enumerate(taps()):
        Y += Xp[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx, :] @ W[t].T
    if b is not None:
        Y += b
    return Y


def conv_input_grad(G, W):
    """dL/dX of Y = conv(X, W) given G = dL/dY."""
    nz, ny, nx, _ = G.shape
    Gp = np.pad(G, ((1, 1), (1, 1), (1, 1), (0, 0)))
    dX = np.zeros((nz, ny, nx, W.shape[2]))
    for t, (dx, dy, dz) in enumerate(taps()):
        dX += Gp[1 - dz:1 - dz + nz, 1 - dy:1 - dy + ny, 1 - dx:1 - dx + nx, :] @ W[t]
    return dX


def conv_weight_grad(X, G):
    nz, ny, nx, _ = X.shape
    Xp = np.pad(X, ((1, 1), (1, 1), (1, 1), (0, 0)))
    dW = np.zeros((27, G.shape[3], X.shape[3]))
    g2 = G.reshape(-1, G.shape[3])
    for t, (dx, dy, dz) in enumerate(taps()):
        dW[t] = g2.T @ Xp[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx, :].reshape(-1, X.shape[3])
    return dW


def _counts(shape):
    """children per parent of a (nz, ny, nx) box, on the half box."""
    c = [np.where(2 * np.arange((n + 1) // 2) + 1 < n, 2, 1) for n in shape]
    return c[0][:, None, None] * c[1][None, :, None] * c[2][None, None, :]


def _pad_even(X):
    nz, ny, nx = X.shape[:3]
    return np.pad(X, ((0, nz % 2), (0, ny % 2), (0, nx % 2), (0, 0)))


def pool(X):
    Xe = _pad_even(X)
    nz, ny, nx, c = Xe.shape
    s = Xe.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).sum(axis=(1, 3, 5))
    return s / _counts(X.shape[:3])[..., None]


def pool_t(H, fine_shape):
    """transpose of pool onto a box of fine_shape (nz, ny, nx)."""
    Hs = H / _counts(fine_shape)[..., None]
    return upsample(Hs, fine_shape)


def upsample(X, fine_shape):
    nz, ny, nx = fine_shape
    return X[np.arange(nz) // 2][:, np.arange(ny) // 2][:, :, np.arange(nx) // 2]


def upsample_t(H, coarse_shape):
    He = _pad_even(H)
    nz, ny, nx, c = He.shape
    return He.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).sum(axis=(1, 3, 5))[:coarse_shape[0], :coarse_shape[1],
                                                                              :coarse_shape[2]]


def _in16(x4):
    nz, ny, nx, _ = x4.shape
    X = np.zeros((nz, ny, nx, 16))
    X[..., :4] = x4
    return X


def forward(theta, x4, emulate=False):
    """Returns (Y[..., :3], cache)."""
    P = unpack(theta)
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    Wt = {k: (r(w), b) for k, (w, b) in P.items()}
    relu = lambda a: np.maximum(a, 0.0)
    c = {}
    F = r(_in16(np.asarray(x4, np.float64)))
    c["F"] = F
    c["a0"] = r(relu(conv(F, *Wt["E0A"])))
    c["s0"] = r(relu(conv(c["a0"], *Wt["E0B"])))
    c["q1"] = r(pool(c["s0"]))
    c["a1"] = r(relu(conv(c["q1"], *Wt["E1A"])))
    c["s1"] = r(relu(conv(c["a1"], *Wt["E1B"])))
    c["q2"] = r(pool(c["s1"]))
    c["a2"] = r(relu(conv(c["q2"], *Wt["E2A"])))
    c["b2"] = r(relu(conv(c["a2"], *Wt["E2B"])))
    c["u1"] = upsample(c["b2"], c["s1"].shape[:3])
    c["c1"] = r(relu(conv(c["u1"], *Wt["D1A"]) + c["s1"]))
    c["t1"] = r(relu(conv(c["c1"], *Wt["D1B"])))
    c["u0"] = upsample(c["t1"], c["s0"].shape[:3])
    c["c0"] = r(relu(conv(c["u0"], *Wt["D0A"]) + c["s0"]))
    c["t0"] = r(relu(conv(c["c0"], *Wt["D0B"])))
    c["Y"] = conv(c["t0"], *Wt["HEAD"])
    c["W"] = Wt
    return c["Y"][..., :3], c


def loss_and_grad(theta, x4, target, emulate=False):
    """(loss, gradient vector) of mean over positions and 3 channels of (net(x) - target)^2."""
    y, c = forward(theta, x4, emulate)
    Wt = c["W"]
    r = bf16 if emulate else (lambda a: a)
    n = y.size
    d = y - target
    loss = float((d * d).sum() / n)
    gY = np.zeros(c["Y"].shape)
    gY[..., :3] = 2.0 * d / n
    gY = r(gY)
    grads = {}

    def wb(name, X, G):
        grads[name] = (conv_weight_grad(X, G), G.sum(axis=(0, 1, 2)))

    mask = lambda g, a: np.where(a > 0, g, 0.0)
    wb("HEAD", c["t0"], gY)
    g_t0 = r(mask(conv_input_grad(gY, Wt["HEAD"][0]), c["t0"]))
    wb("D0B", c["c0"], g_t0)
    g_c0 = r(mask(conv_input_grad(g_t0, Wt["D0B"][0]), c["c0"]))
    wb("D0A", c["u0"], g_c0)
    h_u0 = conv_input_grad(g_c0, Wt["D0A"][0])
    g_t1 = r(mask(upsample_t(h_u0, c["t1"].shape[:3]), c["t1"]))
    wb("D1B", c["c1"], g_t1)
    g_c1 = r(mask(conv_input_grad(g_t1, Wt["D1B"][0]), c["c1"]))
    wb("D1A", c["u1"], g_c1)
    h_u1 = conv_input_grad(g_c1, Wt["D1A"][0])
    g_b2 = r(mask(upsample_t(h_u1, c["b2"].shape[:3]), c["b2"]))
    wb("E2B", c["a2"], g_b2)
    g_a2 = r(mask(conv_input_grad(g_b2, Wt["E2B"][0]), c["a2"]))
    wb("E2A", c["q2"], g_a2)
    h_q2 = conv_input_grad(g_a2, Wt["E2A"][0])
    g_s1 = r(mask(pool_t(h_q2, c["s1"].shape[:3]) + g_c1, c["s1"]))
    wb("E1B", c["a1"], g_s1)
    g_a1 = r(mask(conv_input_grad(g_s1, Wt["E1B"][0]), c["a1"]))
    wb("E1A", c["q1"], g_a1)
    h_q1 = conv_input_grad(g_a1, Wt["E1A"][0])
    g_s0 = r(mask(pool_t(h_q1, c["s0"].shape[:3]) + g_c0, c["s0"]))
    wb("E0B", c["a0"], g_s0)
    g_a0 = r(mask(conv_input_grad(g_s0, Wt["E0B"][0]), c["a0"]))
    wb("E0A", c["F"], g_a0)
    vec = np.concatenate([np.concatenate([grads[k][0].ravel(), grads[k][1]]) for k in LAYERS])
    return loss, vec


def adam_step(w, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma and Ba 2015, Algorithm 1) in float32, the
    arithmetic order of the GPU kernel: m = b1 m + (1 - b1) g,
    v = b2 v + (1 - b2) g^2, w -= lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps).
    t is the 1-based step count. Returns new (w, m, v)."""
    f = np.float32
    w, m, v, g = (np.asarray(a, np.float32) for a in (w, m, v, g))
    m = f(b1) * m + f(1 - f(b1)) * g
    v = f(b2) * v + f(1 - f(b2)) * g * g
    c1 = f(1) - np.power(f(b1), f(t))
    c2 = f(1) - np.power(f(b2), f(t))
    w = w - f(lr) * (m / c1) / (np.sqrt(v / c2) + f(eps))
    return w, m, v
