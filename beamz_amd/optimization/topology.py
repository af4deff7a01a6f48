"""Topology-optimisation helpers (beamz/optimization/topology.py:22-306, examples/4_topo.py:52-131).

Density filters, projections and the masked update used by the adjoint loop.
The heavy parts of that loop (forward and adjoint FDTD, gradient
accumulation) run on the GPU through ``beamz_amd.optimization.adjoint``; these
host helpers operate on the small design-region arrays.
"""
from __future__ import annotations

from typing import Callable, Optional, Tuple

import numpy as np
from numpy.lib.stride_tricks import sliding_window_view

from .optimizers import Optimizer


def _window_sums(a, r):
    return sliding_window_view(a, (2 * r + 1, 2 * r + 1)).sum(axis=(-2, -1))


def _masked_box_blur(values, mask, radius):
    """Edge-padded box mean over the masked neighbourhood (topology.py:31-48)."""
    r = int(max(0, round(radius)))
    if r < 1:
        return values
    v = np.pad(values, r, mode="edge")
    m = np.pad(mask.astype(float), r, mode="constant", constant_values=0.0)
    num = sliding_window_view(v, (2 * r + 1, 2 * r + 1))
    wm = sliding_window_view(m, (2 * r + 1, 2 * r + 1))
    s = (num * wm).sum(axis=(-2, -1))
    w = wm.sum(axis=(-2, -1))
    return s / np.where(w == 0.0, 1.0, w)


def masked_box_blur(values, mask, radius):
    """(blurred, weights) of the example's filter (4_topo.py:52-76)."""
    r = int(max(0, radius))
    mv = np.where(mask, values, 0.0)
    if r <= 0:
        return mv, np.ones_like(mv, dtype=float)
    v = np.pad(mv, r, mode="edge")
    m = np.pad(mask.astype(float), r, mode="constant", constant_values=0.0)
    wm = sliding_window_view(m, (2 * r + 1, 2 * r + 1))
    s = (sliding_window_view(v, (2 * r + 1, 2 * r + 1)) * wm).sum(axis=(-2, -1))
    w = wm.sum(axis=(-2, -1))
    w = np.where(w == 0.0, 1.0, w)
    return np.where(mask, s / w, 0.0), np.where(mask, w, 1.0)


def masked_box_blur_backprop(grad_output, mask, weights, radius):
    """Adjoint of ``masked_box_blur`` w.r.t. its masked input (4_topo.py:79-95)."""
    r = int(max(0, radius))
    g = np.where(mask, grad_output, 0.0) / np.where(weights == 0.0, 1.0, weights)
    if r <= 0:
        return np.where(mask, g, 0.0)
    return np.where(mask, _window_sums(np.pad(g, r), r), 0.0)


def smoothed_heaviside(value, beta, eta):
    """tanh projection and its derivative (4_topo.py:107-117)."""
    beta = max(beta, 1e-6)
    den = np.tanh(beta * eta) + np.tanh(beta * (1.0 - eta))
    if abs(den) < 1e-9:
        den = 1e-9 if den >= 0 else -1e-9
    th = np.tanh(beta * (value - eta))
    return (np.tanh(beta * eta) + th) / den, (beta * (1.0 - th ** 2)) / den


def apply_density_filters(values, mask, radius, beta, eta=0.5):
    """Masked blur then projection: (blurred, projected, weights, derivative) (4_topo.py:120-131)."""
    b, w = masked_box_blur(values, mask, radius)
    b = np.clip(np.where(mask, b, 0.0), 0.0, 1.0)
    p, d = smoothed_heaviside(b, beta, eta)
    p = np.clip(np.where(mask, p, 0.0), 0.0, 1.0)
    d = np.where(mask, d, 0.0)
    d = np.where((p <= 0.0) | (p >= 1.0), 0.0, d)
    return b, p, w, d


def cone_filter(x, radius):
    """Normalised linear-hat (cone) filter, weights max(0, r - |d|), zero padding."""
    r = int(radius)
    yy, xx = np.mgrid[-r:r + 1, -r:r + 1]
    K = np.maximum(0.0, float(r) - np.hypot(xx, yy))
    num = (sliding_window_view(np.pad(x, r), K.shape) * K).sum(axis=(-2, -1))
    den = (sliding_window_view(np.pad(np.ones_like(x), r), K.shape) * K).sum(axis=(-2, -1))
    return num / den


def cone_filter_backprop(g, radius):
    r = int(radius)
    yy, xx = np.mgrid[-r:r + 1, -r:r + 1]
    K = np.maximum(0.0, float(r) - np.hypot(xx, yy))
    den = (sliding_window_view(np.pad(np.ones_like(g), r), K.shape) * K).sum(axis=(-2, -1))
    return (sliding_window_view(np.pad(g / den, r), K.shape) * K[::-1, ::-1]).sum(axis=(-2, -1))


def apply_density_update(density, gradient, mask, *, learning_rate, blur_radius=0,
                         blur_fn: Optional[Callable] = None) -> Tuple[np.ndarray, np.ndarray, np.ndarray, float, float]:
    """Masked, clipped gradient step (topology.py:51-115)."""
    if density.shape != gradient.shape:
        raise ValueError("Gradient and density must share shape.")
    if mask.shape != density.shape:
        raise ValueError("Mask must share the same shape as density.")
    g = np.where(mask, gradient, 0.0).astype(float, copy=False)
    if blur_fn is not None:
        g = blur_fn(g, mask)
    elif blur_radius > 0:
        g = _masked_box_blur(g, mask, blur_radius)
    g = np.where(mask, g, 0.0)
    new = np.where(mask, np.clip(density + learning_rate * g, 0.0, 1.0), density)
    delta = np.where(mask, new - density, 0.0)
    if np.any(mask):
        return new, g, delta, float(np.linalg.norm(delta[mask])), float(np.max(np.abs(delta[mask])))
    return new, g, delta, 0.0, 0.0


def blur_density(density, radius, *, backend="numpy"):
    """Separable box blur, 'same' convolution (topology.py:138-150)."""
    r = max(0, int(round(radius)))
    if r <= 0:
        return density
    k = np.ones(2 * r + 1) / (2 * r + 1)
    out = np.apply_along_axis(lambda row: np.convolve(row, k, mode="same"), 1, density)
    return np.apply_along_axis(lambda col: np.convolve(col, k, mode="same"), 0, out)


def project_density(density, beta, eta):
    """Logistic projection (topology.py:153-161)."""
    if beta <= 0:
        return density
    return 1.0 / (1.0 + np.exp(-beta * (density - eta)))


def compute_overlap_gradient(forward_fields, adjoint_fields, *, monitor=None):
    """Re(conj(adjoint) * forward) of the last Ez snapshots (topology.py:164-178)."""
    f = _extract_field(forward_fields, "Ez")
    a = _extract_field(adjoint_fields, "Ez")
    if f is None or a is None:
        raise ValueError("Forward/adjoint data must contain an 'Ez' entry.")
    g = np.real(np.conj(a) * f)
    return g[tuple(slice(0, min(p, q)) for p, q in zip(f.shape, a.shape))]


def apply_optimizer_step(density, gradient, optimizer_state, *, method="adam", learning_rate=1e-2, **kwargs):
    if gradient.shape != density.shape:
        raise ValueError("Gradient and density must share shape.")
    opt = Optimizer(method=method, learning_rate=learning_rate, options=kwargs) if optimizer_state is None \
        else optimizer_state["optimizer"]
    return np.clip(density + opt.step(gradient), 0.0, 1.0), {"optimizer": opt}


def _extract_field(container, key):
    if container is None:
        return None
    if isinstance(container, dict):
        v = container.get(key)
        return v[-1] if isinstance(v, list) and v else v
    return getattr(container, key, None)
