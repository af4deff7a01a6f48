"""FP64 restatement of the reference topology-optimisation helpers.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows:
  * masked box blur / backprop ........ examples/4_topo.py:52-95
  * smoothed Heaviside projection ..... examples/4_topo.py:107-117
  * density filter chain .............. examples/4_topo.py:120-131
  * reference overlap gradient ........ examples/4_topo.py:279-286 (Q26 heuristic)
  * apply_density_update .............. beamz/optimization/topology.py:31-115
  * Adam .............................. beamz/optimization/optimizers.py:23-63
  * cone filter (config "radius-3 cone filter"): no reference counterpart; the
    standard linear-hat density filter w = max(0, r - |d|), normalised.

Pinned by: the reference's own KATs (tests/test_topology_updates.py:7-55,
re-expressed in tests/test_oracle_topo.py), closed forms of Adam's first step
(update = -lr * sign(g)), tanh projection identities (P(eta)=eta-symmetric,
P(0)=0, P(1)=1), and adjointness <blur(x), y> = <x, blur^T(y)> checked by
brute force.
"""
from __future__ import annotations

import numpy as np
from numpy.lib.stride_tricks import sliding_window_view


def masked_box_blur(values, mask, radius):
    """4_topo.py:52-76 -> (blurred, weights)."""
    radius = int(max(0, radius))
    mv = np.where(mask, values, 0.0)
    if radius <= 0:
        return mv, np.where(mask, 1.0, 1.0)
    pv = np.pad(mv, radius, mode="edge")
    pm = np.pad(mask.astype(float), radius, mode="constant", constant_values=0.0)
    win = (2 * radius + 1, 2 * radius + 1)
    vv = sliding_window_view(pv, win)
    mvw = sliding_window_view(pm, win)
    ws = (vv * mvw).sum(axis=(-2, -1))
    w = mvw.sum(axis=(-2, -1))
    w = np.where(w == 0.0, 1.0, w)
    b = ws / w
    b = np.where(mask, b, 0.0)
    w = np.where(mask, w, 1.0)
    return b, w


def masked_box_blur_backprop(g, mask, weights, radius):
    """4_topo.py:79-95."""
    radius = int(max(0, radius))
    g = np.where(mask, g, 0.0)
    weights = np.where(weights == 0.0, 1.0, weights)
    c = g / weights
    if radius <= 0:
        return np.where(mask, c, 0.0)
    p = np.pad(c, radius, mode="constant", constant_values=0.0)
    win = (2 * radius + 1, 2 * radius + 1)
    gi = sliding_window_view(p, win).sum(axis=(-2, -1))
    return np.where(mask, gi, 0.0)


def smoothed_heaviside(v, beta, eta):
    """4_topo.py:107-117 -> (projected, derivative)."""
    beta = max(beta, 1e-6)
    den = np.tanh(beta * eta) + np.tanh(beta * (1.0 - eta))
    if abs(den) < 1e-9:
        den = 1e-9 if den >= 0 else -1e-9
    num = np.tanh(beta * eta) + np.tanh(beta * (v - eta))
    return num / den, (beta * (1.0 - np.tanh(beta * (v - eta)) ** 2)) / den


def apply_density_filters(values, mask, radius, beta, eta=0.5):
    """4_topo.py:120-131 -> (blurred, projected, weights, derivative)."""
    b, w = masked_box_blur(values, mask, radius)
    b = np.where(mask, b, 0.0)
    b = np.clip(b, 0.0, 1.0)
    p, d = smoothed_heaviside(b, beta, eta)
    p = np.where(mask, p, 0.0)
    p = np.clip(p, 0.0, 1.0)
    d = np.where(mask, d, 0.0)
    d = np.where((p <= 0.0) | (p >= 1.0), 0.0, d)
    return b, p, w, d


def cone_kernel(radius):
    r = int(radius)
    yy, xx = np.mgrid[-r:r + 1, -r:r + 1]
    w = np.maximum(0.0, float(r) - np.sqrt(xx * xx + yy * yy))
    return w


def cone_filter(x, radius):
    """Normalised cone (linear-hat) filter with zero padding: y = (K*x) / (K*1)."""
    K = cone_kernel(radius)
    r = K.shape[0] // 2
    xp = np.pad(x, r)
    op = np.pad(np.ones_like(x), r)
    num = (sliding_window_view(xp, K.shape) * K).sum(axis=(-2, -1))
    den = (sliding_window_view(op, K.shape) * K).sum(axis=(-2, -1))
    return num / den


def cone_filter_T(g, radius):
    """Transpose of cone_filter."""
    K = cone_kernel(radius)
    r = K.shape[0] // 2
    op = np.pad(np.ones_like(g), r)
    den = (sliding_window_view(op, K.shape) * K).sum(axis=(-2, -1))
    gp = np.pad(g / den, r)
    return (sliding_window_view(gp, K.shape) * K[::-1, ::-1]).sum(axis=(-2, -1))


def tanh_projection(x, beta, eta=0.5):
    return smoothed_heaviside(x, beta, eta)


def adam_update(grad, lr, state, beta1=0.9, beta2=0.999, epsilon=1e-8):
    """optimizers.py:23-63."""
    if "m" not in state:
        state["m"] = np.zeros_like(grad)
    if "v" not in state:
        state["v"] = np.zeros_like(grad)
    state["t"] = state.get("t", 0) + 1
    m, v, t = state["m"], state["v"], state["t"]
    m *= beta1
    m += (1.0 - beta1) * grad
    v *= beta2
    v += (1.0 - beta2) * (grad ** 2)
    mh = m / (1.0 - beta1 ** t)
    vh = v / (1.0 - beta2 ** t)
    return -lr * mh / (np.sqrt(vh) + epsilon)


def box_blur_topology(values, mask, radius):
    """topology.py:31-48."""
    radius = int(max(0, round(radius)))
    if radius < 1:
        return values
    pv = np.pad(values, radius, mode="edge")
    pm = np.pad(mask.astype(float), radius, mode="constant", constant_values=0.0)
    win = (2 * radius + 1, 2 * radius + 1)
    ws = (sliding_window_view(pv, win) * sliding_window_view(pm, win)).sum(axis=(-2, -1))
    w = sliding_window_view(pm, win).sum(axis=(-2, -1))
    w = np.where(w == 0.0, 1.0, w)
    return ws / w


def apply_density_update(density, gradient, mask, learning_rate, blur_radius=0):
    """topology.py:51-115."""
    if density.shape != gradient.shape:
        raise ValueError("Gradient and density must share shape.")
    if mask.shape != density.shape:
        raise ValueError("Mask must share the same shape as density.")
    mg = np.where(mask, gradient, 0.0).astype(float, copy=False)
    sm = box_blur_topology(mg, mask, blur_radius) if blur_radius > 0 else mg
    sm = np.where(mask, sm, 0.0)
    nd = np.clip(density + learning_rate * sm, 0.0, 1.0)
    nd = np.where(mask, nd, density)
    dd = np.where(mask, nd - density, 0.0)
    if np.any(mask):
        cn = float(np.linalg.norm(dd[mask])); mu = float(np.max(np.abs(dd[mask])))
    else:
        cn = mu = 0.0
    return nd, sm, dd, cn, mu


def overlap_gradient(adj_series, fwd_series):
    """4_topo.py:282-285: grad = sum_k Re(Ez_adj(k) * conj(Ez_fwd(N-1-k)))."""
    g = np.zeros_like(np.real(fwd_series[0]), dtype=np.float64)
    fw = list(fwd_series)
    for a in adj_series:
        if not fw:
            break
        g += np.real(a * np.conj(fw.pop()))
    return g
