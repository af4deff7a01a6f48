"""FDTD parameter helpers (beamz/helpers.py:21-87) and quiet console stand-ins.

The reference prints rich tables and progress bars (beamz/helpers.py:89-209);
those are cosmetic and kept here as call-compatible functions that only log
through ``logging`` so that the time loop never formats console output.
"""
from __future__ import annotations

import datetime
import logging
from contextlib import contextmanager

import numpy as np

from .const import LIGHT_SPEED

log = logging.getLogger("beamz_amd")


def get_si_scale_and_label(value):
    """Scale factor and unit label for a length (beamz/helpers.py:14-19)."""
    if value >= 1e-3:
        return 1e3, "mm"
    if value >= 1e-6:
        return 1e6, "µm"
    if value >= 1e-9:
        return 1e9, "nm"
    return 1e12, "pm"


def check_fdtd_stability(dt, dx, dy=None, dz=None, n_max=1.0, safety_factor=1.0):
    """Courant check c*dt/min(d) <= safety/sqrt(dims) (beamz/helpers.py:21-53).

    Returns (is_stable, courant_number, limit).
    """
    dims, spacing = 1, dx
    if dy is not None:
        dims, spacing = 2, min(dx, dy)
    if dz is not None:
        dims, spacing = 3, min(dx, dy, dz)
    courant = LIGHT_SPEED * dt / spacing
    limit = safety_factor * (1.0 / np.sqrt(dims))
    return courant <= limit, courant, limit


def calc_optimal_fdtd_params(wavelength, n_max, dims=2, safety_factor=0.999, points_per_wavelength=10):
    """(resolution, dt) from points per wavelength in the densest medium (beamz/helpers.py:55-87)."""
    resolution = wavelength / n_max / points_per_wavelength
    dt = safety_factor * (resolution / (LIGHT_SPEED * np.sqrt(dims)))
    ok, courant, limit = check_fdtd_stability(dt, resolution, dy=resolution if dims >= 2 else None,
                                              dz=resolution if dims >= 3 else None, n_max=n_max)
    if not courant <= limit + 1e-15:
        raise AssertionError("Internal error: calculated time step exceeds stability limit")
    return resolution, dt


def display_status(message, status="info"):
    """Log a status line; "warning" and "error" go out at those levels, everything else at INFO."""
    level = {"warning": logging.WARNING, "error": logging.ERROR}.get(status, logging.INFO)
    log.log(level, "[%s] %s", status, message)


def display_header(title, subtitle=None):
    log.info("%s%s", title, f" - {subtitle}" if subtitle else "")


def display_parameters(params, title="Parameters"):
    log.info("%s: %s", title, params)


def display_results(results, title="Results"):
    log.info("%s: %s", title, results)


def display_simulation_status(progress, metrics=None):
    log.info("progress %.1f%% %s", progress, metrics or "")


def display_optimization_progress(iteration, total, best_value, parameters=None):
    log.info("iteration %d/%d best %s", iteration, total, best_value)


def display_time_elapsed(start_time):
    if start_time is not None:
        log.info("elapsed %s", datetime.datetime.now() - start_time)


def tree_view(data, title="Structure"):
    log.info("%s: %s", title, data)


def code_preview(code, language="python"):
    log.info("%s", code)


class _NullProgress:
    def add_task(self, *a, **k):
        return 0

    def update(self, *a, **k):
        pass


@contextmanager
def create_rich_progress():
    yield _NullProgress()
