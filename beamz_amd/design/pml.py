"""Conductivity-graded absorbing boundary (beamz/design/pml.py:5-76).

``get_conductivity`` evaluates the reference's profile: a quadratic onset to
0.01 over the first 5% of the layer depth, then ((d-0.05)/0.95)**order, scaled
by sigma_max = sigma_factor * 1.2 / (eta * dx) with eta = sqrt(mu0/(eps0*eps))
and, when a time step is supplied, by 1/(1 + alpha_max*d**2). The native
rasterizer evaluates the same expression in C++ with the same operation order.
"""
from __future__ import annotations


import numpy as np

from ..const import EPS_0, MU_0


class PML:
    def __init__(self, region_type, position, size, orientation, polynomial_order=2.0, sigma_factor=1.0,
                 alpha_max=0.1):
        self.region_type = region_type
        self.position = position
        self.orientation = orientation
        self.polynomial_order = polynomial_order
        self.sigma_factor = sigma_factor
        self.alpha_max = alpha_max
        if region_type == "rect":
            self.width, self.height = size
        else:
            self.radius = size

    def get_profile(self, normalized_distance):
        d = min(max(normalized_distance, 0.0), 1.0)
        if d < 0.05:
            s = 0.01 * (d / 0.05) ** 2
        else:
            s = ((d - 0.05) / 0.95) ** self.polynomial_order
        return s, self.alpha_max * d ** 2

    def _depth(self, x, y):
        px, py = self.position
        if self.region_type == "rect":
            if not (px <= x <= px + self.width and py <= y <= py + self.height):
                return None
            o = self.orientation
            if o == "left":
                return 1.0 - (x - px) / self.width
            if o == "right":
                return (x - px) / self.width
            if o == "top":
                return (y - py) / self.height
            if o == "bottom":
                return 1.0 - (y - py) / self.height
            return None
        r = np.hypot(x - px, y - py)
        if r > self.radius:
            return None
        ox, oy = x - px, y - py
        blocked = {"top-left": ox > 0 or oy < 0, "top-right": ox < 0 or oy < 0,
                   "bottom-left": ox > 0 or oy > 0, "bottom-right": ox < 0 or oy > 0}
        if blocked.get(self.orientation, False):
            return None
        return r / self.radius

    def get_conductivity(self, x, y, dx=None, dt=None, eps_avg=None):
        if dx is not None and eps_avg is not None:
            eta = np.sqrt(MU_0 / (EPS_0 * eps_avg))
            smax = 1.2 / (eta * dx)
            smax *= self.sigma_factor
        else:
            smax = 1.0
        d = self._depth(x, y)
        if d is None:
            return 0.0
        s, a = self.get_profile(d)
        c = smax * s
        if dt is not None:
            c *= 1.0 / (1.0 + a)
        return c

    def copy(self):
        size = (self.width, self.height) if self.region_type == "rect" else self.radius
        return PML(self.region_type, self.position, size, self.orientation, self.polynomial_order,
                   self.sigma_factor, self.alpha_max)
