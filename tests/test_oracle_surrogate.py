"""The surrogate oracle's layers, pinned by brute force, scipy and finite differences."""
import numpy as np
from scipy.signal import correlate2d

from oracle import surrogate_ref as S


def test_conv3x3_brute_force_and_scipy():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 5, 6, 3)); w = rng.standard_normal((4, 3, 3, 3)); b = rng.standard_normal(4)
    y = S.conv3x3(x, w, b)
    yb = np.zeros_like(y)
    for bb in range(2):
        for h in range(5):
            for ww in range(6):
                for o in range(4):
                    acc = b[o]
                    for dy in range(3):
                        for dx in range(3):
                            hh, xx = h + dy - 1, ww + dx - 1
                            if 0 <= hh < 5 and 0 <= xx < 6:
                                acc += x[bb, hh, xx, :] @ w[o, dy, dx, :]
                    yb[bb, h, ww, o] = acc
    assert np.allclose(y, yb, rtol=1e-12, atol=1e-12)
    ys = sum(correlate2d(x[1, :, :, i], w[2, :, :, i], mode="same") for i in range(3)) + b[2]
    assert np.allclose(y[1, :, :, 2], ys, rtol=1e-12, atol=1e-12)


def test_conv3x3_backward_finite_differences():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 4, 5, 2)); w = rng.standard_normal((3, 3, 3, 2)); gy = rng.standard_normal((1, 4, 5, 3))
    dx, dw, db = S.conv3x3_backward(x, w, gy)
    f = lambda x_, w_: float((S.conv3x3(x_, w_) * gy).sum())
    h = 1e-6
    for idx in [(0, 1, 2, 1), (0, 0, 0, 0), (0, 3, 4, 1)]:
        e = np.zeros_like(x); e[idx] = h
        assert abs((f(x + e, w) - f(x - e, w)) / (2 * h) - dx[idx]) < 1e-7
    for idx in [(1, 0, 2, 1), (2, 1, 1, 0)]:
        e = np.zeros_like(w); e[idx] = h
        assert abs((f(x, w + e) - f(x, w - e)) / (2 * h) - dw[idx]) < 1e-7
    assert np.allclose(db, gy.sum(axis=(0, 1, 2)))


def test_pool_closed_forms_and_bf16():
    x = np.arange(16.0).reshape(1, 4, 4, 1)
    assert np.allclose(S.avgpool2(x)[0, :, :, 0], [[2.5, 4.5], [10.5, 12.5]])
    assert np.allclose(S.avgpool2_backward(np.ones((1, 2, 2, 1))), 0.25)
    assert S.bf16(1.0 + 2 ** -8) == 1.0                     # halfway between 1 and 1+2^-7 rounds to even
    assert S.bf16(1.0 + 3 * 2 ** -9) == 1.0 + 2 ** -7
    assert S.bf16(-3.0) == -3.0


def test_network_gradient_finite_differences():
    P = S.init_params(3)
    rng = np.random.default_rng(4)
    x = rng.random((2, 16, 16))
    gout = rng.standard_normal((2, 8))
    out, cache = S.forward(P, x)
    G = S.backward(P, x, cache, gout)
    f = lambda P_, x_: float((S.forward(P_, x_)[0] * gout).sum())
    h = 1e-6
    for name, idx in (("w00", (3, 1, 2, 0)), ("w21", (5, 0, 1, 7)), ("b30", (11,)), ("fw", (2, 40))):
        Pp = {k: v.copy() for k, v in P.items()}; Pp[name][idx] += h
        Pm = {k: v.copy() for k, v in P.items()}; Pm[name][idx] -= h
        fd = (f(Pp, x) - f(Pm, x)) / (2 * h)
        assert abs(fd - G[name][idx]) <= 1e-6 * max(1.0, abs(fd)), (name, fd, G[name][idx])
    for idx in ((0, 3, 4), (1, 10, 12)):
        e = np.zeros_like(x); e[idx] = h
        fd = (f(P, x + e) - f(P, x - e)) / (2 * h)
        assert abs(fd - G["x"][idx]) <= 1e-6 * max(1.0, abs(fd)), (idx, fd, G["x"][idx])
