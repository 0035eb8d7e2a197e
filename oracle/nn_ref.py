"""CPU reference of the warm-start network (test infrastructure only, like the rest of
oracle/): numpy restatement of the 3D convolutional encoder-decoder that
gp5_amd/nn.py runs on the GPU, with its backward pass, for gradient checks.

The network has no counterpart in the reference (lucaparisi91/GP5 has no learned
part); this file is the CPU oracle of the kernels in include/gp5_nn.h.

Layout: channels-last arrays [B, D, D, D, C]; weights W[27, Cin, Cout] with
tap = 9 (dz+1) + 3 (dy+1) + (dx+1); 'same' zero padding.
"""
from __future__ import annotations

import numpy as np

LEAK = 0.1


def layer_shapes(c0: int = 16, widths=(16, 32, 64, 128), cout: int = 16):
    w1, w2, w3, w4 = widths
    return [("c1", c0, w1), ("c2", w1, w2), ("c3", w2, w3), ("c4", w3, w4),
            ("c5", w4, w3), ("c6", w3, w2), ("c7", w2, w1), ("c8", w1, cout)]


def taps():
    return [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def shift(X, dz, dy, dx):
    """S[p] = X[p + (dz, dy, dx)] with zeros outside."""
    D = X.shape[1]
    S = np.zeros_like(X)
    S[:, max(0, -dz):D - max(0, dz), max(0, -dy):D - max(0, dy), max(0, -dx):D - max(0, dx)] = \
        X[:, max(0, dz):D - max(0, -dz), max(0, dy):D - max(0, -dy), max(0, dx):D - max(0, -dx)]
    return S


def conv(X, W, b=None):
    Y = np.zeros(X.shape[:4] + (W.shape[2],), dtype=np.result_type(X, W))
    for t, (dz, dy, dx) in enumerate(taps()):
        Y += shift(X, dz, dy, dx) @ W[t]
    if b is not None:
        Y += b
    return Y


def conv_dgrad(dY, W):
    """dX[p] = sum_t dY[p - off(t)] W[t]^T."""
    dX = np.zeros(dY.shape[:4] + (W.shape[1],), dtype=np.result_type(dY, W))
    for t, (dz, dy, dx) in enumerate(taps()):
        dX += shift(dY, -dz, -dy, -dx) @ W[t].T
    return dX


def conv_wgrad(X, dY):
    W = np.zeros((27, X.shape[4], dY.shape[4]), dtype=np.result_type(X, dY))
    dYf = dY.reshape(-1, dY.shape[4])
    for t, (dz, dy, dx) in enumerate(taps()):
        Xf = shift(X, dz, dy, dx).reshape(-1, X.shape[4])
        W[t] = Xf.T @ dYf
    return W


def pool2(X):
    B, D = X.shape[0], X.shape[1]
    return X.reshape(B, D // 2, 2, D // 2, 2, D // 2, 2, -1).mean(axis=(2, 4, 6))


def up2(X):
    return X.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)


def lrelu(z):
    return np.where(z > 0, z, LEAK * z)


def lrelu_grad(a):
    return np.where(a > 0, 1.0, LEAK)


def init(shapes, seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    p = {}
    for name, ci, co in shapes:
        p[name + ".W"] = rng.standard_normal((27, ci, co)) * scale * np.sqrt(2.0 / (27 * ci))
        p[name + ".b"] = rng.standard_normal(co) * 0.01
    return p


def forward(p, X, base):
    s = {}
    s["e1"] = lrelu(conv(X, p["c1.W"], p["c1.b"]))
    s["p1"] = pool2(s["e1"])
    s["e2"] = lrelu(conv(s["p1"], p["c2.W"], p["c2.b"]))
    s["p2"] = pool2(s["e2"])
    s["e3"] = lrelu(conv(s["p2"], p["c3.W"], p["c3.b"]))
    s["p3"] = pool2(s["e3"])
    s["e4"] = lrelu(conv(s["p3"], p["c4.W"], p["c4.b"]))
    s["u4"] = up2(s["e4"])
    s["a5"] = lrelu(conv(s["u4"], p["c5.W"], p["c5.b"]))
    s["d3"] = s["a5"] + s["e3"]
    s["u3"] = up2(s["d3"])
    s["a6"] = lrelu(conv(s["u3"], p["c6.W"], p["c6.b"]))
    s["d2"] = s["a6"] + s["e2"]
    s["u2"] = up2(s["d2"])
    s["a7"] = lrelu(conv(s["u2"], p["c7.W"], p["c7.b"]))
    s["d1"] = s["a7"] + s["e1"]
    s["out"] = conv(s["d1"], p["c8.W"], p["c8.b"])
    y = base + s["out"][..., 0]
    return y, s


def loss_and_grads(p, X, base, target):
    """MSE over all positions and its gradients (float64 backward of forward())."""
    y, s = forward(p, X, base)
    P = y.size
    L = float(np.mean((y - target) ** 2))
    g = {}
    dout = np.zeros_like(s["out"])
    dout[..., 0] = 2.0 * (y - target) / P
    g["c8.W"] = conv_wgrad(s["d1"], dout)
    g["c8.b"] = dout.sum(axis=(0, 1, 2, 3))
    dd1 = conv_dgrad(dout, p["c8.W"])
    de1 = dd1.copy()
    dz7 = dd1 * lrelu_grad(s["a7"])
    g["c7.W"] = conv_wgrad(s["u2"], dz7)
    g["c7.b"] = dz7.sum(axis=(0, 1, 2, 3))
    dd2 = pool2(conv_dgrad(dz7, p["c7.W"])) * 8.0
    de2 = dd2.copy()
    dz6 = dd2 * lrelu_grad(s["a6"])
    g["c6.W"] = conv_wgrad(s["u3"], dz6)
    g["c6.b"] = dz6.sum(axis=(0, 1, 2, 3))
    dd3 = pool2(conv_dgrad(dz6, p["c6.W"])) * 8.0
    de3 = dd3.copy()
    dz5 = dd3 * lrelu_grad(s["a5"])
    g["c5.W"] = conv_wgrad(s["u4"], dz5)
    g["c5.b"] = dz5.sum(axis=(0, 1, 2, 3))
    de4 = pool2(conv_dgrad(dz5, p["c5.W"])) * 8.0
    dz4 = de4 * lrelu_grad(s["e4"])
    g["c4.W"] = conv_wgrad(s["p3"], dz4)
    g["c4.b"] = dz4.sum(axis=(0, 1, 2, 3))
    de3 = de3 + up2(conv_dgrad(dz4, p["c4.W"])) / 8.0
    dz3 = de3 * lrelu_grad(s["e3"])
    g["c3.W"] = conv_wgrad(s["p2"], dz3)
    g["c3.b"] = dz3.sum(axis=(0, 1, 2, 3))
    de2 = de2 + up2(conv_dgrad(dz3, p["c3.W"])) / 8.0
    dz2 = de2 * lrelu_grad(s["e2"])
    g["c2.W"] = conv_wgrad(s["p1"], dz2)
    g["c2.b"] = dz2.sum(axis=(0, 1, 2, 3))
    de1 = de1 + up2(conv_dgrad(dz2, p["c2.W"])) / 8.0
    dz1 = de1 * lrelu_grad(s["e1"])
    g["c1.W"] = conv_wgrad(X, dz1)
    g["c1.b"] = dz1.sum(axis=(0, 1, 2, 3))
    return L, g, y
