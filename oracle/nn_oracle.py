"""CPU oracle of the surrogate network's operations -- TEST INFRASTRUCTURE ONLY.

float64 numpy restatements of the kernels of csrc/surrogate.hip: 3x3
convolution (NHWC, weights [tap][cout][cin], tap = 3*(dy+1)+(dx+1), zero
padding), its weight and data gradients, 2x2 average pooling, nearest 2x
upsampling, the masked MSE loss and an Adam step, plus the whole
encoder-decoder of Eikonal/surrogate.py.  The gradients are pinned by finite
differences of the forward restatements (tests/test_oracle_nn.py).
"""
import numpy as np


def conv3x3(x, w, relu=False):
    N, H, W, Ci = x.shape
    Co = w.shape[1]
    xp = np.zeros((N, H + 2, W + 2, Ci))
    xp[:, 1:-1, 1:-1] = x
    y = np.zeros((N, H, W, Co))
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        y += np.einsum('nhwc,oc->nhwo', xp[:, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W], w[tap])
    return np.maximum(y, 0) if relu else y


def conv3x3_wgrad(x, gy):
    N, H, W, Ci = x.shape
    xp = np.zeros((N, H + 2, W + 2, Ci))
    xp[:, 1:-1, 1:-1] = x
    dw = np.zeros((9, gy.shape[-1], Ci))
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        dw[tap] = np.einsum('nhwo,nhwc->oc', gy, xp[:, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W])
    return dw


def conv3x3_dgrad(gy, w):
    """Gradient w.r.t. the input: the convolution of gy with the weights
    reversed in space and transposed."""
    wt = np.transpose(w[::-1], (0, 2, 1))
    return conv3x3(gy, wt)


def pool2(x):
    N, H, W, C = x.shape
    return x.reshape(N, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def pool2_bwd(g):
    return np.repeat(np.repeat(g, 2, axis=1), 2, axis=2) / 4


def up2(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up2_bwd(g):
    N, H, W, C = g.shape
    return g.reshape(N, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


def mse(pred, target, mask=None):
    """Loss mean(m (p0 - t)^2) on channel 0 and its gradient."""
    m = np.ones_like(target) if mask is None else mask
    d = pred[..., 0] - target
    count = d.size
    g = np.zeros_like(pred)
    g[..., 0] = 2 * m * d / count
    return float(np.sum(m * d * d) / count), g


def adam(w, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return w, m, v


# ---------------------------------------------------------------- network
def network_forward(x, layers):
    """layers: 15 weight arrays in the order enc0a, enc0b, ..., enc3b,
    dec2a, dec2b, dec1a, dec1b, dec0a, dec0b, head (as Surrogate.layers)."""
    c = {}
    h = x
    skips = []
    k = 0
    for lvl in range(4):
        if lvl > 0:
            h = pool2(h)
        c[f'e{lvl}_x'] = h
        a = conv3x3(h, layers[k], True); k += 1
        c[f'e{lvl}_a'] = a
        h = conv3x3(a, layers[k], True); k += 1
        c[f'e{lvl}_b'] = h
        skips.append(h)
    for lvl in (2, 1, 0):
        u = up2(h)
        c[f'd{lvl}_u'] = u
        a = conv3x3(u, layers[k], True); k += 1
        c[f'd{lvl}_a'] = a
        sm = a + skips[lvl]
        c[f'd{lvl}_s'] = sm
        h = conv3x3(sm, layers[k], True); k += 1
        c[f'd{lvl}_b'] = h
    c['h_x'] = h
    return conv3x3(h, layers[k]), c


def network_backward(gy, layers, c):
    """Weight gradients of network_forward (same order as layers)."""
    grads = [None] * len(layers)
    grads[14] = conv3x3_wgrad(c['h_x'], gy)
    g = conv3x3_dgrad(gy, layers[14])
    gskip = {}
    k = {2: 8, 1: 10, 0: 12}
    for lvl in (0, 1, 2):
        g = g * (c[f'd{lvl}_b'] > 0)
        grads[k[lvl] + 1] = conv3x3_wgrad(c[f'd{lvl}_s'], g)
        g = conv3x3_dgrad(g, layers[k[lvl] + 1])
        gskip[lvl] = g.copy()
        g = g * (c[f'd{lvl}_a'] > 0)
        grads[k[lvl]] = conv3x3_wgrad(c[f'd{lvl}_u'], g)
        g = conv3x3_dgrad(g, layers[k[lvl]])
        g = up2_bwd(g)
    for lvl in (3, 2, 1, 0):
        if lvl in gskip:
            g = g + gskip[lvl]
        g = g * (c[f'e{lvl}_b'] > 0)
        grads[2 * lvl + 1] = conv3x3_wgrad(c[f'e{lvl}_a'], g)
        g = conv3x3_dgrad(g, layers[2 * lvl + 1])
        g = g * (c[f'e{lvl}_a'] > 0)
        grads[2 * lvl] = conv3x3_wgrad(c[f'e{lvl}_x'], g)
        if lvl > 0:
            g = conv3x3_dgrad(g, layers[2 * lvl])
            g = pool2_bwd(g)
    return grads
