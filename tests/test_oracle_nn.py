"""Pins the CPU oracle of the surrogate network: gradients of every
operation and of the whole encoder-decoder against central finite
differences of the forward restatements."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), '..', 'oracle'))
import nn_oracle as NN  # noqa: E402


def fd_check(f, x, g_analytic, rng, n=12, eps=1e-6):
    idx = [tuple(rng.integers(0, s) for s in x.shape) for _ in range(n)]
    for i in idx:
        xp, xm = x.copy(), x.copy()
        xp[i] += eps
        xm[i] -= eps
        num = (f(xp) - f(xm)) / (2 * eps)
        assert num == pytest.approx(g_analytic[i], rel=1e-5, abs=1e-8)


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 5, 6, 3))
    w = rng.normal(size=(9, 4, 3))
    y = NN.conv3x3(x, w)
    n, i, j, o = 1, 0, 5, 2
    ref = sum(w[3 * (dy + 1) + dx + 1, o] @ x[n, i + dy, j + dx]
              for dy in (-1, 0, 1) for dx in (-1, 0, 1)
              if 0 <= i + dy < 5 and 0 <= j + dx < 6)
    assert y[n, i, j, o] == pytest.approx(ref)


def test_conv_gradients():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 6, 5, 3))
    w = rng.normal(size=(9, 4, 3))
    r = rng.normal(size=(2, 6, 5, 4))
    fx = lambda xx: np.sum(r * NN.conv3x3(xx, w))
    fw = lambda ww: np.sum(r * NN.conv3x3(x, ww))
    fd_check(fx, x, NN.conv3x3_dgrad(r, w), rng)
    fd_check(fw, w, NN.conv3x3_wgrad(x, r), rng)


def test_pool_up_adjoints():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(2, 8, 6, 3))
    g = rng.normal(size=(2, 4, 3, 3))
    assert np.sum(NN.pool2(x) * g) == pytest.approx(np.sum(x * NN.pool2_bwd(g)))
    y = rng.normal(size=(2, 4, 3, 3))
    G = rng.normal(size=(2, 8, 6, 3))
    assert np.sum(NN.up2(y) * G) == pytest.approx(np.sum(y * NN.up2_bwd(G)))


def test_network_gradients():
    rng = np.random.default_rng(3)
    widths = (4, 4, 8, 8)
    shapes = [(6, 4), (4, 4), (4, 4), (4, 4), (4, 8), (8, 8), (8, 8), (8, 8),
              (8, 8), (8, 8), (8, 4), (4, 4), (4, 4), (4, 4), (4, 1)]
    layers = [rng.normal(scale=0.4, size=(9, co, ci)) for ci, co in shapes]
    x = rng.normal(size=(1, 16, 16, 6))
    t = rng.normal(size=(1, 16, 16))

    def loss(ls):
        y, _ = NN.network_forward(x, ls)
        return NN.mse(y, t)[0]

    y, c = NN.network_forward(x, layers)
    _, gy = NN.mse(y, t)
    grads = NN.network_backward(gy, layers, c)
    for k in (0, 5, 9, 14):
        def f(wk, k=k):
            ls = list(layers)
            ls[k] = wk
            return loss(ls)
        fd_check(f, layers[k], grads[k], rng, n=6)


def test_adam_matches_closed_form_first_step():
    w, g = np.array([1.0, -2.0]), np.array([0.5, -0.25])
    w1, m, v = NN.adam(w, g, np.zeros(2), np.zeros(2), 1, lr=0.1)
    np.testing.assert_allclose(w1, w - 0.1 * np.sign(g), rtol=1e-6)
