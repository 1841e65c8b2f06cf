"""TEST INFRASTRUCTURE ONLY -- FP64 numpy restatement of the learned
correction network (3-D U-Net, 3 levels, 16/32/64 channels, 3x3x3 convs).
Its input is the pair (x, d): an early SART iterate and its last update.

The network has no counterpart in the reference: it is an extension that
predicts a late SART iterate from an early one.  This module is
the mathematical definition the HIP kernels are checked against (forward
activations and every parameter gradient); its own gradients are pinned by
central finite differences in tests/test_unet_oracle.py.  Only tests/ may
import it.

Layout: activations [D][H][W][C] (channels last), conv weights [Cout][27][Cin]
with tap t = (dz * 3 + dy) * 3 + dx reading input voxel (z + dz - 1,
y + dy - 1, x + dx - 1), zero outside the volume; biases [Cout].
"""
from __future__ import annotations

import numpy as np

# (name, input channels, output channels); inputs of c4a / c5a are
# concat(upsample(lower level), skip) in that channel order.
LAYERS = [
    ("c1a", 2, 16), ("c1b", 16, 16),
    ("c2a", 16, 32), ("c2b", 32, 32),
    ("c3a", 32, 64), ("c3b", 64, 64),
    ("c4a", 96, 32), ("c4b", 32, 32),
    ("c5a", 48, 16), ("c5b", 16, 16),
]
HEAD_CHANNELS = 16


def parameter_count():
    n = sum(co * 27 * ci + co for _, ci, co in LAYERS)
    return n + HEAD_CHANNELS + 1


def unpack(flat):
    """Flat parameter vector -> {name: (W [Cout][27][Cin], b [Cout])}, head (w [16], b)."""
    flat = np.asarray(flat, np.float64)
    out, at = {}, 0
    for name, ci, co in LAYERS:
        w = flat[at:at + co * 27 * ci].reshape(co, 27, ci)
        at += co * 27 * ci
        b = flat[at:at + co]
        at += co
        out[name] = (w, b)
    hw = flat[at:at + HEAD_CHANNELS]
    hb = flat[at + HEAD_CHANNELS]
    return out, (hw, hb)


def init_parameters(seed):
    """He-normal weights, zero biases; head weights small so the initial
    correction is near zero."""
    rng = np.random.default_rng(seed)
    parts = []
    for _, ci, co in LAYERS:
        parts.append(rng.normal(0.0, np.sqrt(2.0 / (27 * ci)), co * 27 * ci))
        parts.append(np.zeros(co))
    parts.append(rng.normal(0.0, 1e-3, HEAD_CHANNELS))
    parts.append(np.zeros(1))
    return np.concatenate(parts).astype(np.float32)


def _taps():
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                yield (dz * 3 + dy) * 3 + dx, dz, dy, dx


def conv3(x, w, b):
    """3x3x3 convolution, zero padding 1: y[p, co] = b[co] + sum_t,ci x[p + d(t), ci] w[co, t, ci]."""
    D, H, W, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((D, H, W, w.shape[0])) + b
    for t, dz, dy, dx in _taps():
        y += xp[dz:dz + D, dy:dy + H, dx:dx + W, :] @ w[:, t, :].T
    return y


def conv3_backward(x, w, gy):
    """Gradients of conv3 with respect to x, w and b."""
    D, H, W, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    gxp = np.zeros_like(xp)
    gw = np.zeros_like(w)
    g2 = gy.reshape(-1, gy.shape[-1])
    for t, dz, dy, dx in _taps():
        xs = xp[dz:dz + D, dy:dy + H, dx:dx + W, :].reshape(-1, x.shape[-1])
        gw[:, t, :] = g2.T @ xs
        gxp[dz:dz + D, dy:dy + H, dx:dx + W, :] += gy @ w[:, t, :]
    return gxp[1:-1, 1:-1, 1:-1, :], gw, g2.sum(axis=0)


def pool2(x):
    D, H, W, C = x.shape
    return x.reshape(D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(1, 3, 5))


def pool2_backward(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 0), 2, 1), 2, 2) / 8.0


def up2(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up2_backward(g):
    D, H, W, C = g.shape
    return g.reshape(D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(1, 3, 5))


def relu(x):
    return np.maximum(x, 0.0)


def forward(params, x, d=None):
    """x, d: [D][H][W] (D, H, W multiples of 4; d None reads as zeros).
    Returns (prediction x + delta, cache)."""
    P, (hw, hb) = unpack(params)
    x = np.asarray(x, np.float64)
    d = np.zeros_like(x) if d is None else np.asarray(d, np.float64)
    c = {"x": np.stack([x, d], axis=-1)}

    def layer(name, inp):
        c[name + ".in"] = inp
        y = relu(conv3(inp, *P[name]))
        c[name] = y
        return y

    e1 = layer("c1b", layer("c1a", c["x"]))
    e2 = layer("c2b", layer("c2a", pool2(e1)))
    e3 = layer("c3b", layer("c3a", pool2(e2)))
    d2 = layer("c4b", layer("c4a", np.concatenate([up2(e3), e2], axis=-1)))
    d1 = layer("c5b", layer("c5a", np.concatenate([up2(d2), e1], axis=-1)))
    delta = d1 @ hw + hb
    return c["x"][..., 0] + delta, c


def loss_and_gradient(params, x, target, d=None):
    """L = mean((prediction - target)^2) and dL/dparams (flat, FP64)."""
    P, (hw, hb) = unpack(params)
    pred, c = forward(params, x, d)
    r = pred - np.asarray(target, np.float64)
    L = float(np.mean(r * r))
    g = 2.0 * r / r.size
    grads = {}
    ghw = np.tensordot(c["c5b"], g, axes=([0, 1, 2], [0, 1, 2]))
    ghb = g.sum()
    gd1 = g[..., None] * hw

    def back(name, gy):
        gy = gy * (c[name] > 0)
        gx, gw, gb = conv3_backward(c[name + ".in"], P[name][0], gy)
        grads[name] = (gw, gb)
        return gx

    g5 = back("c5a", back("c5b", gd1))
    gd2, ge1 = up2_backward(g5[..., :32]), g5[..., 32:]
    g4 = back("c4a", back("c4b", gd2))
    ge3, ge2 = up2_backward(g4[..., :64]), g4[..., 64:]
    ge2 = ge2 + pool2_backward(back("c3a", back("c3b", ge3)))
    ge1 = ge1 + pool2_backward(back("c2a", back("c2b", ge2)))
    back("c1a", back("c1b", ge1))
    parts = []
    for name, _, _ in LAYERS:
        gw, gb = grads[name]
        parts += [gw.ravel(), gb]
    parts += [ghw, np.array([ghb])]
    return L, np.concatenate(parts)
