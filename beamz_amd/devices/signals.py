"""Source time signals (beamz/devices/signals.py:5-24).

Same formulas as the reference, including ``gaussian_pulse`` multiplying the
amplitude in twice (envelope and carrier), so it peaks at amplitude**2.
"""
from __future__ import annotations

import numpy as np


def cosine(t, amplitude, frequency, phase):
    return amplitude * np.cos(2 * np.pi * frequency * t + phase)


def sigmoid(t, duration=1, min=0, max=1, t0=0):
    return min + (max - min) * (1 / (1 + np.exp(-10 * (t - duration / 2 - t0) / duration)))


def ramped_cosine(t, amplitude, frequency, phase=None, ramp_duration=None, t_max=None):
    phase = 0 if phase is None else phase
    if ramp_duration is None:
        ramp_duration = t_max / 10
    env_on = sigmoid(t, min=0, max=1, duration=ramp_duration, t0=0)
    env_on *= cosine(t, amplitude=amplitude, frequency=frequency, phase=phase)
    env_on *= sigmoid(t, min=1, max=0, duration=ramp_duration, t0=t_max - ramp_duration)
    return env_on


def gaussian(t, amplitude, center, width):
    return amplitude * np.exp(-(t - center) ** 2 / (2 * width ** 2))


def gaussian_pulse(t, amplitude, center, width, frequency, phase):
    return gaussian(t, amplitude, center, width) * cosine(t, amplitude, frequency, phase)


def plot_signal(signals, t):
    """Plotting is outside this package's scope; kept as a no-op for API compatibility."""
    return None
