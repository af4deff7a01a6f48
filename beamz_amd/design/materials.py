"""Materials (beamz/design/materials.py:2-217).

``Material`` is a dispersionless medium. ``CustomMaterial`` has spatially varying
permittivity, permeability and conductivity, each given by a function of
(x, y) or (x, y, z) or by a 2D array on a rectangle (``bounds``) read through
``scipy.interpolate.RegularGridInterpolator``, with the reference's
constructor, getters, ``update_grid`` and ``copy``. Its getters follow the
reference's calling convention: ``get_permittivity(x, y)`` calls a function
with two arguments, ``get_permittivity(x, y, z)`` with three (the rasterizer
passes z = 0, so a two-argument function fails there and the rasterizer takes
the defaults, as the reference's does), and a grid is read at (y, x).
"""
from __future__ import annotations

import numpy as np

_PROPS = ("permittivity", "permeability", "conductivity")
_DEFAULTS = {"permittivity": 1.0, "permeability": 1.0, "conductivity": 0.0}


class Material:
    def __init__(self, permittivity=1.0, permeability=1.0, conductivity=0.0):
        self.permittivity = permittivity
        self.permeability = permeability
        self.conductivity = conductivity

    def get_sample(self):
        return self.permittivity, self.permeability, self.conductivity

    def copy(self):
        return Material(self.permittivity, self.permeability, self.conductivity)


class CustomMaterial:
    """Spatially varying material for inverse design (beamz/design/materials.py:14-217)."""

    def __init__(self, permittivity_func=None, permeability_func=None, conductivity_func=None,
                 permittivity_grid=None, permeability_grid=None, conductivity_grid=None,
                 bounds=None, interpolation="linear"):
        self.permittivity_func = permittivity_func
        self.permeability_func = permeability_func
        self.conductivity_func = conductivity_func
        self.permittivity_grid = permittivity_grid
        self.permeability_grid = permeability_grid
        self.conductivity_grid = conductivity_grid
        self.bounds = bounds
        self.interpolation = interpolation
        self.default_permittivity = _DEFAULTS["permittivity"]
        self.default_permeability = _DEFAULTS["permeability"]
        self.default_conductivity = _DEFAULTS["conductivity"]
        for name in _PROPS:
            self._interpolate(name)

    def _interpolate(self, name):
        """A grid over ``bounds`` = ((x0, x1), (y0, y1)), rows along y, read at (y, x); outside the bounds
        the property's default."""
        grid = getattr(self, name + "_grid")
        if grid is None or self.bounds is None:
            return
        from scipy.interpolate import RegularGridInterpolator
        g = np.asarray(grid)
        (x0, x1), (y0, y1) = self.bounds
        axes = (np.linspace(y0, y1, g.shape[0]), np.linspace(x0, x1, g.shape[1]))
        setattr(self, "_" + name + "_interpolator",
                RegularGridInterpolator(axes, g, method=self.interpolation, bounds_error=False,
                                        fill_value=getattr(self, "default_" + name)))

    def _value(self, name, x, y, z):
        f = getattr(self, name + "_func")
        if f is not None:
            return f(x, y) if z is None else f(x, y, z)
        interp = getattr(self, "_" + name + "_interpolator", None)
        if interp is not None:
            return interp(np.column_stack([np.atleast_1d(y), np.atleast_1d(x)]))
        return getattr(self, "default_" + name)

    def _label(self, name):
        g = getattr(self, name + "_grid")
        if g is not None:
            return f"grid({np.min(g):.3f}-{np.max(g):.3f})"
        if getattr(self, name + "_func") is not None:
            return "function"
        return getattr(self, "default_" + name)

    permittivity = property(lambda self: self._label("permittivity"))
    permeability = property(lambda self: self._label("permeability"))
    conductivity = property(lambda self: self._label("conductivity"))

    def get_permittivity(self, x, y, z=None):
        return self._value("permittivity", x, y, z)

    def get_permeability(self, x, y, z=None):
        return self._value("permeability", x, y, z)

    def get_conductivity(self, x, y, z=None):
        return self._value("conductivity", x, y, z)

    def get_sample(self, x=0, y=0, z=None):
        return self.get_permittivity(x, y, z), self.get_permeability(x, y, z), self.get_conductivity(x, y, z)

    def update_grid(self, property_name, new_grid):
        """Replace one property's grid (an optimisation step) and rebuild its interpolator."""
        if property_name not in _PROPS:
            raise ValueError(f"Unknown property: {property_name}")
        setattr(self, property_name + "_grid", new_grid)
        self._interpolate(property_name)

    def copy(self):
        """Functions are shared; grids are copied."""
        grids = {name + "_grid": (None if getattr(self, name + "_grid") is None
                                  else np.array(getattr(self, name + "_grid"), copy=True)) for name in _PROPS}
        return CustomMaterial(permittivity_func=self.permittivity_func, permeability_func=self.permeability_func,
                              conductivity_func=self.conductivity_func, bounds=self.bounds,
                              interpolation=self.interpolation, **grids)
