"""Design container (beamz/design/core.py:15-90, 125-147, 399-443, 559-607).

``structures[0]`` is the background rectangle; later structures paint over
earlier ones. With ``auto_pml`` the reference's eight PML regions (four edge
rectangles, four corner quarter-discs) and four display-only rectangles
(``is_pml=True``, skipped by the rasterizer) are appended.

GDS export and import (core.py:445-553) go through this package's own GDSII
reader and writer (``design/gds.py``; the reference uses gdspy). Polygon union
(core.py:159-200) needs shapely, which is not available: ``unify_polygons``
reports that and returns False, as the reference does without shapely.
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

    def unify_polygons(self):
        """core.py:159-200. Merging touching polygons of one material needs shapely, which is absent here: as
        the reference does without it, report and return False, leaving the structures as they are."""
        display_status("Shapely library is required for polygon unification. Please install with: pip install shapely",
                       "error")
        return False

    def export_gds(self, output_file):
        """core.py:477-553: the structures (not the PML display regions, sources or monitors) as GDSII boundaries,
        one layer per material in order of first appearance (permittivity, permeability, conductivity), in
        microns with a 1 nm database unit. A polygon with interiors is written as one keyhole boundary."""
        from ..devices.monitors import Monitor
        from ..devices.sources import GaussianSource, ModeSource
        from .gds import write_gds
        self.unify_polygons()
        groups = {}
        for s in self.structures:
            if getattr(s, "is_pml", False) or isinstance(s, (ModeSource, GaussianSource, Monitor)):
                continue
            m = getattr(s, "material", None)
            if m is None:
                continue
            key = (getattr(m, "permittivity", 1.0), getattr(m, "permeability", 1.0), getattr(m, "conductivity", 0.0))
            groups.setdefault(key, []).append(s)
        scale = 1e6
        polys = []
        for layer, members in enumerate(groups.values()):
            for s in members:
                verts = getattr(s, "vertices", None)
                if not isinstance(s, Polygon) or not verts:
                    continue
                outline = [(v[0] * scale, v[1] * scale) for v in verts]
                holes = [[(v[0] * scale, v[1] * scale) for v in h] for h in (getattr(s, "interiors", None) or [])]
                polys.append((layer, 0, outline, holes))
        write_gds(output_file, polys, unit=1e-6, precision=1e-9)
        print(f"GDS file saved as '{output_file}' with {len(groups)} material-based layers")
        for layer, key in enumerate(groups):
            print(f"Layer {layer}: εᵣ={key[0]:.1f}, μᵣ={key[1]:.1f}, σ={key[2]:.2e} S/m")
        display_status(f"Created design with size: {self.width:.2e} x {self.height:.2e} x {self.depth:.2e} m")

    @staticmethod
    def import_gds(gds_file, default_depth=1e-6):
        """core.py:445-475: every structure's polygons (references flattened) by layer, as Polygons of the given
        depth in metres, in ``layers[layer]`` and in ``structures`` of a new default Design."""
        from .gds import read_gds
        unit, cells, skipped = read_gds(gds_file)
        design = Design()
        n = 0
        for specs in cells.values():
            for (layer, _datatype), polys in specs.items():
                design.layers.setdefault(layer, [])
                for pts in polys:
                    poly = Polygon(vertices=[(float(x) * unit, float(y) * unit) for x, y in pts], depth=default_depth)
                    design.layers[layer].append(poly)
                    design.structures.append(poly)
                    n += 1
        if default_depth > 0:
            design.is_3d = True
            design.depth = default_depth
        print(f"Imported {n} polygons from '{gds_file}' into Design object.")
        if skipped:
            display_status(f"{skipped} PATH/TEXT elements of '{gds_file}' were not imported", "warning")
        if design.is_3d:
            print(f"3D design with depth: {design.depth:.2e} m")
        return design

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
