"""Design container (beamz/design/core.py:15-90, 125-147, 399-443, 559-607).

``structures[0]`` is the background rectangle; later structures paint over
earlier ones. With ``auto_pml`` the reference's eight PML regions (four edge
rectangles, four corner quarter-discs) and four display-only rectangles
(``is_pml=True``, skipped by the rasterizer) are appended. Polygon union and
GDS import/export are outside the FDTD hot path and are not provided.
"""
from __future__ import annotations

import numpy as np

from ..const import µm
from ..helpers import display_status
from .materials import Material
from .pml import PML
from .structures import Polygon, Rectangle


class Design:
    def __init__(self, width=5 * µm, height=5 * µm, depth=0, material=None, color=None, border_color="black",
                 auto_pml=True, pml_size=None):
        if material is None:
            material = Material(permittivity=1.0, permeability=1.0, conductivity=0.0)
        self.structures = [Rectangle(position=(0, 0, 0), width=width, height=height, depth=depth,
                                     material=material, color=color)]
        self.sources = []
        self.monitors = []
        self.boundaries = []
        self.width = width
        self.height = height
        self.depth = depth
        self.border_color = border_color
        self.time = 0
        self.layers = {}
        self.is_3d = not (depth is None or depth == 0)
        if auto_pml:
            self._init_boundaries(pml_size)

    def __str__(self):
        return f"Design with {len(self.structures)} structures ({'3D' if self.is_3d else '2D'})"

    def __iadd__(self, structure):
        self.add(structure)
        return self

    def _init_boundaries(self, pml_size=None):
        """core.py:42-90: auto thickness from the densest material seen so far."""
        if pml_size is None:
            emax = 1.0
            for s in self.structures:
                m = getattr(s, "material", None)
                if m is not None and hasattr(m, "permittivity"):
                    emax = max(emax, m.permittivity)
            lam = 1.55e-6 / np.sqrt(emax)
            lo = 1.5 * lam
            hi = min(self.width, self.height) * 0.3
            pml_size = max(lo, min(hi, min(self.width, self.height) / 3))
            display_status(f"Auto-selected PML size: {pml_size:.2e} m", "info")
        W, H = self.width, self.height
        self.boundaries += [
            PML("rect", (0, 0), (pml_size, H), "left"),
            PML("rect", (W - pml_size, 0), (pml_size, H), "right"),
            PML("rect", (0, H - pml_size), (W, pml_size), "top"),
            PML("rect", (0, 0), (W, pml_size), "bottom"),
            PML("corner", (0, 0), pml_size, "bottom-left"),
            PML("corner", (W, 0), pml_size, "bottom-right"),
            PML("corner", (0, H), pml_size, "top-left"),
            PML("corner", (W, H), pml_size, "top-right"),
        ]
        shade = Material(permittivity=1.0, permeability=1.0, conductivity=0.0)
        for pos, w, h in (((0, 0), pml_size, H), ((W - pml_size, 0), pml_size, H), ((0, 0), W, pml_size),
                          ((0, H - pml_size), W, pml_size)):
            self.structures.append(Rectangle(position=pos, width=w, height=h, material=shade, color="none",
                                             is_pml=True))

    def add(self, structure):
        """core.py:125-146: devices go to their lists and to the paint order."""
        from ..devices.monitors import Monitor
        from ..devices.sources import GaussianSource, ModeSource
        if isinstance(structure, (ModeSource, GaussianSource)):
            self.sources.append(structure)
        elif isinstance(structure, Monitor):
            self.monitors.append(structure)
        self.structures.append(structure)
        if hasattr(structure, "depth") and structure.depth != 0:
            self.is_3d = True
        elif hasattr(structure, "position") and len(structure.position) > 2 and structure.position[2] != 0:
            self.is_3d = True
        elif getattr(structure, "vertices", None):
            if any(len(v) > 2 and v[2] != 0 for v in structure.vertices):
                self.is_3d = True

    def get_material_value(self, x, y, z=0, dx=None, dt=None):
        """core.py:399-443: topmost structure's material at a point, plus PML conductivity."""
        eps, mu, sig = 1.0, 1.0, 0.0
        # The reference tests isinstance(Polygon) first, so every polygon-derived
        # structure (display-only PML rectangles included) is matched by its outline.
        for s in reversed(self.structures):
            if isinstance(s, Polygon) and s.point_in_polygon(x, y, z):
                eps, mu, sig = s.material.get_sample()
                break
        extra = 0.0
        if dx is not None:
            for b in self.boundaries:
                extra += b.get_conductivity(x, y, dx=dx, dt=dt, eps_avg=eps)
        return [eps, mu, sig + extra]

    def copy(self):
        bg = self.structures[0].material if self.structures and hasattr(self.structures[0], "material") else None
        d = Design(width=self.width, height=self.height, material=bg, auto_pml=False)
        d.structures, d.sources, d.monitors, d.boundaries = [], [], [], []
        from ..devices.monitors import Monitor
        from ..devices.sources import GaussianSource, ModeSource
        for s in self.structures:
            c = s.copy() if hasattr(s, "copy") else s
            if hasattr(c, "design"):
                c.design = d
            d.structures.append(c)
            if isinstance(c, (ModeSource, GaussianSource)):
                d.sources.append(c)
            elif isinstance(c, Monitor):
                d.monitors.append(c)
        d.boundaries = [b.copy() for b in self.boundaries]
        d.is_3d, d.depth, d.border_color, d.time = self.is_3d, self.depth, self.border_color, self.time
        d.layers = dict(self.layers)
        return d
