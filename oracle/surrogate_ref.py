"""FP64 restatement of the surrogate network's layers: the checker of the BF16-MFMA kernels.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

No reference counterpart (the surrogate is new). Layers, NHWC layout:
  conv3x3 (zero padding 1, stride 1): y[b,h,w,o] = bias[o] + sum_{dy,dx,i} x[b,h+dy-1,w+dx-1,i] w[o,dy,dx,i]
  ReLU, 2x2 average pooling, global average pooling, fully connected, MSE loss.
``bf16`` rounds to bfloat16 (round to nearest even), the storage format of the
GPU activations and weights, so the oracle can reproduce what the GPU stores.

Pinned by (tests/test_oracle_surrogate.py): the convolution against brute-force
loops and against scipy.signal.correlate2d per channel pair, every backward
function against central finite differences of its forward, and pooling
against closed forms.
"""
from __future__ import annotations

import numpy as np

LEVELS = ((1, 32), (32, 64), (64, 128), (128, 256))   # (input channels, output channels) of each level


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def conv3x3(x, w, b=None):
    B, H, W, Ci = x.shape
    Co = w.shape[0]
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((B, H, W, Co))
    for dy in range(3):
        for dx in range(3):
            y += xp[:, dy:dy + H, dx:dx + W, :] @ w[:, dy, dx, :].T
    if b is not None:
        y += b
    return y


def conv3x3_backward(x, w, gy):
    """(dx, dw, db) of conv3x3 for the upstream gradient gy."""
    B, H, W, Ci = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    gp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    for dy in range(3):
        for dx in range(3):
            gp[:, dy:dy + H, dx:dx + W, :] += gy @ w[:, dy, dx, :]
            dw[:, dy, dx, :] = np.einsum("bhwo,bhwi->oi", gy, xp[:, dy:dy + H, dx:dx + W, :])
    return gp[:, 1:-1, 1:-1, :], dw, gy.sum(axis=(0, 1, 2))


def avgpool2(x):
    B, H, W, C = x.shape
    return x.reshape(B, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def avgpool2_backward(gy):
    return np.repeat(np.repeat(gy, 2, axis=1), 2, axis=2) / 4.0


def init_params(seed=0):
    """He-normal conv weights [Co,3,3,Ci], zero biases, FC 256 -> 8."""
    rng = np.random.default_rng(seed)
    P = {}
    for lvl, (ci, co) in enumerate(LEVELS):
        for s, (a, o) in enumerate(((ci, co), (co, co))):
            P[f"w{lvl}{s}"] = rng.standard_normal((o, 3, 3, a)) * np.sqrt(2.0 / (9 * a))
            P[f"b{lvl}{s}"] = np.zeros(o)
    P["fw"] = rng.standard_normal((8, 256)) * np.sqrt(1.0 / 256)
    P["fb"] = np.zeros(8)
    return P


def forward(P, x, round_bf16=False):
    """x: (B, H, W) density. Returns (out (B, 8), cache). With round_bf16 every stored activation and the
    weights are rounded to bfloat16 as the GPU stores them."""
    r = bf16 if round_bf16 else (lambda a: a)
    a = r(x[..., None])
    cache = []
    for lvl in range(4):
        for s in range(2):
            w, b = r(P[f"w{lvl}{s}"]), P[f"b{lvl}{s}"]
            z = conv3x3(a, w, b)
            y = r(np.maximum(z, 0.0))
            cache.append((a, w, y))
            a = y
        if lvl < 3:
            a = r(avgpool2(a))
    g = a.mean(axis=(1, 2))
    out = g @ P["fw"].T + P["fb"]
    return out, (cache, g)


def backward(P, x, cache, gout):
    """Gradients of sum(out * gout) w.r.t. every parameter and the input density (FP64, no rounding)."""
    convs, g = cache
    G = {"fw": gout.T @ g, "fb": gout.sum(axis=0)}
    gg = gout @ P["fw"]
    a = convs[-1][2]
    B, H, W, C = a.shape
    ga = np.broadcast_to(gg[:, None, None, :] / (H * W), a.shape).copy()
    idx = len(convs) - 1
    for lvl in range(3, -1, -1):
        if lvl < 3:
            ga = avgpool2_backward(ga)
        for s in (1, 0):
            xin, w, y = convs[idx]
            gz = ga * (y > 0)
            gx, gw, gb = conv3x3_backward(xin, w, gz)
            G[f"w{lvl}{s}"], G[f"b{lvl}{s}"] = gw, gb
            ga = gx
            idx -= 1
    G["x"] = ga[..., 0]
    return G
