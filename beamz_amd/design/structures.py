"""Geometry primitives with the reference's attribute contract.

Mirrors beamz/design/structures.py:6-454: the rasterizer reads ``vertices``
(counter-clockwise exterior, 3-tuples), ``interiors`` (holes, orientation
kept), ``position``, ``width``/``height``/``depth``, ``radius`` and
``inner_radius``/``outer_radius``; those attributes and the even-odd
containment rule are part of parity. Plotting methods are not provided.
"""
from __future__ import annotations

import math

import numpy as np


def _as3(v, z=0.0):
    if len(v) == 2:
        return (v[0], v[1], z)
    if len(v) == 3:
        return (v[0], v[1], v[2])
    raise ValueError(f"Vertex must have 2 or 3 coordinates, got {len(v)}")


def _signed_area2(pts):
    acc = 0
    n = len(pts)
    for a in range(n):
        b = (a + 1) % n
        acc += pts[a][0] * pts[b][1]
        acc -= pts[b][0] * pts[a][1]
    return acc


def _crossings(x, y, path):
    """Even-odd ray test along +x (structures.py:178-196 rule, x/y only)."""
    if not path:
        return False
    n = len(path)
    inside = False
    ax, ay = path[0][0], path[0][1]
    for step in range(1, n + 1):
        bx, by = path[step % n][0], path[step % n][1]
        if y > min(ay, by) and y <= max(ay, by) and x <= max(ax, bx):
            xint = (y - ay) * (bx - ax) / (by - ay) + ax if ay != by else ax
            if ax == bx or x <= xint:
                inside = not inside
        ax, ay = bx, by
    return inside


class Polygon:
    """Closed polygon (optionally with holes) on a z-plane (structures.py:6-207)."""

    def __init__(self, vertices=None, material=None, color=None, optimize=False, interiors=None, depth=0, z=0):
        self.vertices = self._ccw3(vertices if vertices is not None else [], z)
        self.interiors = [[_as3(v, 0.0) for v in hole] for hole in (interiors or []) if hole]
        self.material = material
        self.optimize = optimize
        self.color = color
        self.depth = depth if depth is not None else 0
        self.z = z if z is not None else 0

    @staticmethod
    def _ccw3(vertices, z=0.0):
        if not vertices:
            return []
        v3 = [_as3(v, 0.0) for v in vertices]
        if len(v3) >= 3 and _signed_area2(v3) < 0:
            v3 = v3[::-1]
        return [(x, y, zz) for (x, y, zz) in v3]

    def shift(self, x, y, z=0):
        self.vertices = [(v[0] + x, v[1] + y, v[2] + z) for v in self.vertices]
        self.interiors = [[(v[0] + x, v[1] + y, v[2] + z) for v in h] for h in self.interiors if h]
        return self

    def _centroid(self):
        n = len(self.vertices)
        return (sum(v[0] for v in self.vertices) / n, sum(v[1] for v in self.vertices) / n,
                sum(v[2] for v in self.vertices) / n)

    def scale(self, s_x, s_y=None, s_z=None):
        if s_y is None:
            s_y = s_x
        if s_z is None:
            s_z = 1.0 if s_y != s_x else s_x
        if self.vertices:
            cx, cy, cz = self._centroid()
            f = lambda v: (cx + (v[0] - cx) * s_x, cy + (v[1] - cy) * s_y, cz + (v[2] - cz) * s_z)
            self.vertices = [f(v) for v in self.vertices]
            self.interiors = [[f(v) for v in h] for h in self.interiors if h]
        return self

    def rotate(self, angle, axis="z", point=None):
        if not self.vertices:
            return self
        a = np.radians(angle)
        ca, sa = np.cos(a), np.sin(a)
        if point is None:
            cx, cy, cz = self._centroid()
        else:
            cx, cy, cz = point[0], point[1], (point[2] if len(point) > 2 else 0)
        if axis == "z":
            f = lambda v: (cx + (v[0] - cx) * ca - (v[1] - cy) * sa, cy + (v[0] - cx) * sa + (v[1] - cy) * ca, v[2])
        elif axis == "x":
            f = lambda v: (v[0], cy + (v[1] - cy) * ca - (v[2] - cz) * sa, cz + (v[1] - cy) * sa + (v[2] - cz) * ca)
        elif axis == "y":
            f = lambda v: (cx + (v[0] - cx) * ca + (v[2] - cz) * sa, v[1], cz - (v[0] - cx) * sa + (v[2] - cz) * ca)
        else:
            raise ValueError(f"Invalid rotation axis '{axis}'. Must be 'x', 'y', or 'z'.")
        self.vertices = [f(v) for v in self.vertices]
        self.interiors = [[f(v) for v in h] for h in self.interiors if h]
        return self

    def copy(self):
        return Polygon(vertices=list(self.vertices), interiors=[list(h) for h in self.interiors],
                       material=self.material, color=self.color, optimize=self.optimize,
                       depth=self.depth, z=self.z)

    def get_bounding_box(self):
        if not self.vertices:
            return (0, 0, 0, 0, 0, 0)
        xs = [v[0] for v in self.vertices]
        ys = [v[1] for v in self.vertices]
        zs = [v[2] for v in self.vertices]
        return (min(xs), min(ys), min(zs), max(xs), max(ys), max(zs))

    def point_in_polygon(self, x, y, z=None):
        if not self.vertices or not _crossings(x, y, self.vertices):
            return False
        return not any(h and _crossings(x, y, h) for h in self.interiors)


class Rectangle(Polygon):
    """Axis-aligned box; the rasterizer's fast path keys on ``vertices[0] == position`` (structures.py:209-265)."""

    def __init__(self, position=(0, 0, 0), width=1, height=1, depth=1, material=None, color=None,
                 is_pml=False, optimize=False, z=None):
        if len(position) not in (2, 3):
            raise ValueError("Position must be (x,y) or (x,y,z)")
        if z is not None:
            position = (position[0], position[1], z)
        elif len(position) == 2:
            position = (position[0], position[1], 0.0)
        x, y, zp = position
        corners = [(x, y, zp), (x + width, y, zp), (x + width, y + height, zp), (x, y + height, zp)]
        super().__init__(vertices=corners, material=material, color=color, optimize=optimize, depth=depth, z=zp)
        self.position = position
        self.width = width
        self.height = height
        self.depth = depth
        self.is_pml = is_pml

    def shift(self, x, y, z=0):
        self.position = (self.position[0] + x, self.position[1] + y, self.position[2] + z)
        return super().shift(x, y, z)

    def _refresh_box(self):
        xs = [v[0] for v in self.vertices]; ys = [v[1] for v in self.vertices]; zs = [v[2] for v in self.vertices]
        self.position = (min(xs), min(ys), min(zs))
        self.width = max(xs) - min(xs)
        self.height = max(ys) - min(ys)
        self.depth = max(zs) - min(zs)

    def rotate(self, angle, axis="z", point=None):
        super().rotate(angle, axis, point)
        self._refresh_box()
        return self

    def scale(self, s_x, s_y=None, s_z=None):
        if s_y is None:
            s_y = s_x
        if s_z is None:
            s_z = 1.0 if s_y != s_x else s_x
        super().scale(s_x, s_y, s_z)
        self.width *= s_x
        self.height *= s_y
        self.depth *= s_z
        return self

    def copy(self):
        r = Rectangle(self.position, self.width, self.height, self.depth, self.material, self.color,
                      self.is_pml, self.optimize)
        r.vertices = [tuple(v) for v in self.vertices]
        return r


def _arc(cx, cy, z, r, thetas):
    return [(cx + r * np.cos(t), cy + r * np.sin(t), z) for t in thetas]


class Circle(Polygon):
    """Disc, rasterized analytically in 2D (structures.py:267-297)."""

    def __init__(self, position=(0, 0), radius=1, points=32, material=None, color=None, optimize=False,
                 depth=0, z=0):
        position = _as3(position)
        th = np.linspace(0, 2 * np.pi, points, endpoint=False)
        super().__init__(vertices=_arc(position[0], position[1], position[2], radius, th), material=material,
                         color=color, optimize=optimize, depth=depth, z=z)
        self.position = position
        self.radius = radius
        self.points = points

    def shift(self, x, y, z=0):
        self.position = (self.position[0] + x, self.position[1] + y, self.position[2] + z)
        return super().shift(x, y, z)

    def scale(self, s_x, s_y=None, s_z=None):
        self.radius *= s_x
        th = np.linspace(0, 2 * np.pi, self.points, endpoint=False)
        self.vertices = _arc(self.position[0], self.position[1], self.position[2], self.radius, th)
        return self

    def copy(self):
        return Circle(self.position, self.radius, self.points, self.material, self.color, self.optimize,
                      self.depth, self.z)


class Ring(Polygon):
    """Annulus, rasterized analytically in 2D (structures.py:299-358)."""

    def __init__(self, position=(0, 0), inner_radius=1, outer_radius=2, material=None, color=None,
                 optimize=False, points=256, depth=0, z=None):
        if len(position) not in (2, 3):
            raise ValueError("Position must be (x,y) or (x,y,z)")
        if z is not None:
            position = (position[0], position[1], z)
        position = _as3(position)
        th = np.linspace(0, 2 * np.pi, points, endpoint=False)
        outer = _arc(position[0], position[1], position[2], outer_radius, th)
        inner = _arc(position[0], position[1], position[2], inner_radius, th[::-1])
        super().__init__(vertices=outer, interiors=[inner], material=material, color=color,
                         optimize=optimize, depth=depth, z=position[2])
        self.points = points
        self.position = position
        self.inner_radius = inner_radius
        self.outer_radius = outer_radius

    def shift(self, x, y, z=0):
        self.position = (self.position[0] + x, self.position[1] + y, self.position[2] + z)
        return super().shift(x, y, z)

    def scale(self, s_x, s_y=None, s_z=None):
        self.inner_radius *= s_x
        self.outer_radius *= s_x
        th = np.linspace(0, 2 * np.pi, self.points, endpoint=False)
        p = self.position
        self.vertices = _arc(p[0], p[1], p[2], self.outer_radius, th)
        self.interiors = [_arc(p[0], p[1], p[2], self.inner_radius, th[::-1])]
        return self

    def copy(self):
        return Ring(self.position, self.inner_radius, self.outer_radius, self.material, self.color,
                    self.optimize, self.points, self.depth, self.z)


class CircularBend(Polygon):
    """Arc-shaped waveguide (structures.py:360-419).

    Note: the reference rasterizer routes any object with inner/outer radii
    through its ring path, so in 2D a bend is rasterized as the full annulus
    clipped to the bend's bounding box; this mirror keeps that behaviour.
    """

    def __init__(self, position=(0, 0), inner_radius=1, outer_radius=2, angle=90, rotation=0, material=None,
                 facecolor=None, optimize=False, points=64, depth=0, z=0):
        position = _as3(position)
        self.points = points
        self.position = position
        self.inner_radius = inner_radius
        self.outer_radius = outer_radius
        self.angle = angle
        self.rotation = rotation
        super().__init__(vertices=self._outline(), material=material, color=facecolor, optimize=optimize,
                         depth=depth, z=z)

    def _outline(self):
        th = np.linspace(0, np.radians(self.angle), self.points)
        rot = np.radians(self.rotation)
        p = self.position
        return _arc(p[0], p[1], p[2], self.outer_radius, th + rot) + _arc(p[0], p[1], p[2], self.inner_radius,
                                                                          th[::-1] + rot)

    def shift(self, x, y, z=0):
        self.position = (self.position[0] + x, self.position[1] + y, self.position[2] + z)
        return super().shift(x, y, z)

    def rotate(self, angle, axis="z", point=None):
        if axis == "z":
            self.rotation = (self.rotation + angle) % 360
        return super().rotate(angle, axis, point or self.position)

    def scale(self, s_x, s_y=None, s_z=None):
        self.inner_radius *= s_x
        self.outer_radius *= s_x
        self.vertices = self._outline()
        return self

    def copy(self):
        return CircularBend(self.position, self.inner_radius, self.outer_radius, self.angle, self.rotation,
                            self.material, self.color, self.optimize, self.points, self.depth, self.z)


class Taper(Polygon):
    """Linear taper from input to output width along +x (structures.py:421-454)."""

    def __init__(self, position=(0, 0), input_width=1, output_width=0.5, length=1, material=None, color=None,
                 optimize=False, depth=0, z=0):
        position = _as3(position)
        x, y, zz = position
        verts = [(x, y - input_width / 2, zz), (x + length, y - output_width / 2, zz),
                 (x + length, y + output_width / 2, zz), (x, y + input_width / 2, zz)]
        super().__init__(vertices=verts, material=material, color=color, optimize=optimize, depth=depth, z=zz)
        self.position = position
        self.input_width = input_width
        self.output_width = output_width
        self.length = length

    def rotate(self, angle, axis="z", point=None):
        super().rotate(angle, axis, point)
        xs = [v[0] for v in self.vertices]; ys = [v[1] for v in self.vertices]; zs = [v[2] for v in self.vertices]
        self.position = (min(xs), min(ys), min(zs))
        self.length = max(xs) - min(xs)
        return self

    def copy(self):
        t = Taper(self.position, self.input_width, self.output_width, self.length, self.material, self.color,
                  self.optimize, self.depth, self.z)
        t.vertices = [tuple(v) for v in self.vertices]
        return t
