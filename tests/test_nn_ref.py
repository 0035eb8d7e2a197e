"""The CPU reference of the warm-start network (oracle/nn_ref.py), pinned by exact
identities: the data and weight gradients of the 3x3x3 convolution are its adjoints,
a single-tap kernel is a shift, and the whole backward pass matches central finite
differences in float64 on a small network."""
import numpy as np
import pytest

from oracle import nn_ref as R


def test_conv_identity_and_shift():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((2, 6, 6, 6, 3))
    W = np.zeros((27, 3, 3))
    W[13] = np.eye(3)  # centre tap
    assert np.allclose(R.conv(X, W), X)
    W = np.zeros((27, 3, 3))
    W[14] = np.eye(3)  # tap (0, 0, +1): Y[x] = X[x + 1], zero at the last x
    Y = R.conv(X, W)
    assert np.allclose(Y[:, :, :, :-1], X[:, :, :, 1:])
    assert np.allclose(Y[:, :, :, -1], 0.0)


def test_conv_gradients_are_adjoints():
    rng = np.random.default_rng(2)
    X = rng.standard_normal((2, 5, 5, 5, 3))
    W = rng.standard_normal((27, 3, 4))
    G = rng.standard_normal((2, 5, 5, 5, 4))
    lhs = np.sum(R.conv(X, W) * G)
    assert np.sum(X * R.conv_dgrad(G, W)) == pytest.approx(lhs, rel=1e-12)
    assert np.sum(W * R.conv_wgrad(X, G)) == pytest.approx(lhs, rel=1e-12)


def test_pool_up_adjoint():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((1, 4, 4, 4, 2))
    Z = rng.standard_normal((1, 2, 2, 2, 2))
    assert np.sum(R.pool2(X) * Z) == pytest.approx(np.sum(X * R.up2(Z)) / 8.0, rel=1e-12)


def test_backward_matches_finite_differences():
    shapes = R.layer_shapes(c0=2, widths=(2, 3, 4, 3), cout=2)
    p = R.init(shapes, seed=4)
    rng = np.random.default_rng(5)
    D = 8
    X = rng.standard_normal((1, D, D, D, 2))
    base = rng.standard_normal((1, D, D, D)) * 0.1
    target = rng.standard_normal((1, D, D, D))
    L, g, _ = R.loss_and_grads(p, X, base, target)
    h = 1e-6
    for name in ["c1.W", "c3.W", "c4.b", "c5.W", "c7.b", "c8.W"]:
        flat = p[name].reshape(-1)
        for idx in rng.choice(flat.size, size=min(4, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + h
            Lp = R.loss_and_grads(p, X, base, target)[0]
            flat[idx] = old - h
            Lm = R.loss_and_grads(p, X, base, target)[0]
            flat[idx] = old
            fd = (Lp - Lm) / (2 * h)
            an = g[name].reshape(-1)[idx]
            assert an == pytest.approx(fd, rel=1e-5, abs=1e-10), (name, idx)
