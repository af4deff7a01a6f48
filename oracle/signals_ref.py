"""FP64 restatement of the reference signals and FDTD parameter helpers.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows beamz/devices/signals.py:5-24 and beamz/helpers.py:21-87.
Pinned by closed forms in tests/test_oracle_fdtd.py: the sigmoid midpoint
value 1/2, the A^2 amplitude of gaussian_pulse at its centre (quirk Q11), the
Courant limit 1/sqrt(dims), and dx = lambda/(n*ppw).
"""
from __future__ import annotations

import numpy as np

from .fdtd_ref import LIGHT_SPEED


def cosine(t, amplitude, frequency, phase):
    return amplitude * np.cos(2 * np.pi * frequency * t + phase)


def sigmoid(t, duration=1, min=0, max=1, t0=0):
    return min + (max - min) * (1 / (1 + np.exp(-10 * (t - duration / 2 - t0) / duration)))


def ramped_cosine(t, amplitude, frequency, phase=None, ramp_duration=None, t_max=None):
    if phase is None:
        phase = 0
    if ramp_duration is None:
        ramp_duration = t_max / 10
    s = sigmoid(t, min=0, max=1, duration=ramp_duration, t0=0)
    s *= cosine(t, amplitude=amplitude, frequency=frequency, phase=phase)
    s *= sigmoid(t, min=1, max=0, duration=ramp_duration, t0=t_max - ramp_duration)
    return s


def gaussian(t, amplitude, center, width):
    return amplitude * np.exp(-(t - center) ** 2 / (2 * width ** 2))


def gaussian_pulse(t, amplitude, center, width, frequency, phase):
    return gaussian(t, amplitude, center, width) * cosine(t, amplitude, frequency, phase)


def check_fdtd_stability(dt, dx, dy=None, dz=None, n_max=1.0, safety_factor=1.0):
    dims = 1
    ms = dx
    if dy is not None:
        dims = 2
        ms = min(dx, dy)
    if dz is not None:
        dims = 3
        ms = min(dx, dy, dz)
    courant = LIGHT_SPEED * dt / ms
    lim = safety_factor * (1.0 / np.sqrt(dims))
    return courant <= lim, courant, lim


def calc_optimal_fdtd_params(wavelength, n_max, dims=2, safety_factor=0.999, points_per_wavelength=10):
    res = wavelength / n_max / points_per_wavelength
    dt = safety_factor * (res / (LIGHT_SPEED * np.sqrt(dims)))
    return res, dt
