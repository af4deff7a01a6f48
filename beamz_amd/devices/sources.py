"""Sources (beamz/devices/sources.py:7-483).

``GaussianSource``: soft Gaussian current patch added into Ez after the E
update, weights exp(-(dx^2/(2 wx^2) + dy^2/(2 wy^2))) over +-max(1, round(3w/d))
cells (beamz/simulation/helper.py:56-118; the 2D patch is divided by 4).

``ModeSource``: a line (2D) or plane (3D) of [amplitude, x, y, z] samples
injected at their nearest cells; with ``enforce_direction`` (default True) the
cell one step upstream of each sample is zeroed (helper.py:8-54). Modes come
from ``solve_modes`` (numerical) or ``slab_mode_source`` (analytical).
"""
from __future__ import annotations

import numpy as np

from ..const import LIGHT_SPEED, µm
from .mode import slab_mode_source, solve_modes


def _pos3(p):
    if p is None:
        return None
    if len(p) == 2:
        return (p[0], p[1], 0)
    if len(p) == 3:
        return tuple(p)
    raise ValueError(f"Position must be 2D (x,y) or 3D (x,y,z), got {len(p)} dimensions")


class GaussianSource:
    def __init__(self, position=(0, 0), width=1.0 * µm, signal=0):
        self.position = _pos3(position)
        self.width = width
        self.signal = signal

    @property
    def position_2d(self):
        return (self.position[0], self.position[1])

    def copy(self):
        return GaussianSource(position=self.position, width=self.width, signal=self.signal)


class ModeSource:
    def __init__(self, design, position=None, width=None, height=None, wavelength=1.55 * µm, signal=0, direction="+x",
                 orientation=None, npml=20, num_modes=2, grid_resolution=2000, mode_solver="num_eigen",
                 start=None, end=None):
        if start is not None and end is not None:
            self.start, self.end = _pos3(start), _pos3(end)
            self._position = tuple((a + b) / 2 for a, b in zip(self.start, self.end))
            self.width = float(np.linalg.norm(np.subtract(self.end, self.start)))
            self.height = 0
            self.orientation = orientation
        else:
            if position is None:
                raise ValueError("Either (start, end) or position must be specified")
            self._position = _pos3(position)
            self.width = width if width is not None else wavelength
            self.height = height if height is not None else 0
            if orientation is None:
                orientation = {"+x": "yz", "-x": "yz", "+y": "xz", "-y": "xz", "+z": "xy", "-z": "xy"}.get(direction, "yz")
            if orientation not in ("xy", "xz", "yz"):
                raise ValueError(f"Invalid orientation: {orientation}. Must be 'xy', 'xz', or 'yz'")
            self.orientation = orientation
            x, y, z = self._position
            half = self.width / 2
            if orientation == "yz":
                self.start, self.end = (x, y - half, z), (x, y + half, z)
            else:
                self.start, self.end = (x - half, y, z), (x + half, y, z)
        self.wavelength = wavelength
        self.design = design
        self.signal = signal
        self.direction = direction
        self.npml = npml
        self.num_modes = num_modes
        self.grid_resolution = grid_resolution
        self.mode_solver = mode_solver
        self.enforce_direction = True
        self.dL = wavelength / grid_resolution
        self.omega = 2 * np.pi * LIGHT_SPEED / wavelength
        eps_1d = self.get_eps_1d()
        if mode_solver == "analytical":
            self.effective_indices, self.mode_vectors = self._analytic_modes(eps_1d)
        else:
            self.effective_indices, self.mode_vectors = solve_modes(eps_1d, self.omega, self.dL, npml=npml,
                                                                    m=num_modes)
        if self.height and self.height > 0:
            self.mode_profiles = self._plane_profiles()
        else:
            self.mode_profiles = [self.get_xy_mode_line(self.mode_vectors, q) for q in range(self.mode_vectors.shape[1])]

    @property
    def position(self):
        return self._position

    @property
    def position_2d(self):
        return ((self.start[0] + self.end[0]) / 2, (self.start[1] + self.end[1]) / 2)

    def get_eps_1d(self):
        """Permittivity sampled every dL along the source line (sources.py:294-305)."""
        n = int(np.hypot(self.end[0] - self.start[0], self.end[1] - self.start[1]) / self.dL)
        xs = np.linspace(self.start[0], self.end[0], n)
        ys = np.linspace(self.start[1], self.end[1], n)
        zs = np.linspace(self.start[2], self.end[2], n)
        return np.array([self.design.get_material_value(a, b, c)[0] for a, b, c in zip(xs, ys, zs)], dtype=float)

    def _analytic_modes(self, eps_1d):
        n = eps_1d.size
        x = np.linspace(0, np.hypot(self.end[0] - self.start[0], self.end[1] - self.start[1]), n)
        ncore, nclad = np.sqrt(np.max(eps_1d)), np.sqrt(np.min(eps_1d))
        core = np.where(eps_1d > np.max(eps_1d) * 0.9)[0]
        w = (core[-1] - core[0]) * self.dL if len(core) else self.wavelength
        vecs = np.zeros((n, self.num_modes), dtype=complex)
        neffs = []
        for q in range(self.num_modes):
            try:
                # centred at x0 = 0 (the line start), as sources.py:153-156 calls it
                E, ne = slab_mode_source(x=x, w=w, n_WG=ncore, n0=nclad, wavelength=self.wavelength, ind_m=q)
                vecs[:, q] = E
                neffs.append(ne)
            except Exception:
                neffs.append(0)
        return np.array(neffs), vecs

    def get_xy_mode_line(self, vecs, mode_number):
        n = vecs.shape[0]
        xs = np.linspace(self.start[0], self.end[0], n)
        ys = np.linspace(self.start[1], self.end[1], n)
        zs = np.linspace(self.start[2], self.end[2], n)
        return [[vecs[j, mode_number], xs[j], ys[j], zs[j]] for j in range(n)]

    def _plane_profiles(self):
        """Separable cross-section profile for 3D injection (sources.py:323-376)."""
        x0, y0, z0 = self._position
        ny = max(8, int(round(self.width / self.dL)))
        nz = max(8, int(round(self.height / self.dL)))
        ys = np.linspace(y0 - self.width / 2, y0 + self.width / 2, ny)
        zs = np.linspace(z0 - self.height / 2, z0 + self.height / 2, nz)
        ey = np.array([self.design.get_material_value(x0, y, z0)[0] for y in ys])
        ez = np.array([self.design.get_material_value(x0, y0, z)[0] for z in zs])

        def width_of(c, e):
            idx = np.where(e >= 0.9 * np.max(e))[0]
            return c[idx[-1]] - c[idx[0]] if idx.size else max(c[-1] - c[0], self.dL)

        def prof(c, e, w, ctr):
            try:
                return slab_mode_source(x=c, w=w, n_WG=np.sqrt(e.max()), n0=np.sqrt(e.min()),
                                        wavelength=self.wavelength, ind_m=0, x0=ctr)[0]
            except Exception:
                return np.ones_like(c)

        py = prof(ys, ey, width_of(ys, ey), y0)
        pz = prof(zs, ez, width_of(zs, ez), z0)
        py = py / np.max(np.abs(py) + 1e-12)
        pz = pz / np.max(np.abs(pz) + 1e-12)
        f = np.outer(np.abs(py), np.abs(pz))
        f /= np.max(f) + 1e-12
        return [[[f[a, b], x0, ys[a], zs[b]] for a in range(ny) for b in range(nz)]]

    def copy(self):
        return ModeSource(self.design, start=self.start, end=self.end, wavelength=self.wavelength, signal=self.signal,
                          direction=self.direction, npml=self.npml, num_modes=self.num_modes,
                          grid_resolution=self.grid_resolution, mode_solver=self.mode_solver)
