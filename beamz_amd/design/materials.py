"""Dispersionless materials (beamz/design/materials.py:2-9).

``CustomMaterial`` (spatially varying properties, materials.py:14-217) is
accepted for API compatibility with function-valued properties; the native
rasterizer samples it at cell centres the way the reference's rectangle fast
path does for interior cells.
"""
from __future__ import annotations


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
    """Function-valued material; callables take (x, y[, z]) and return a scalar."""

    def __init__(self, permittivity_func=None, permeability_func=None, conductivity_func=None,
                 permittivity_grid=None, permeability_grid=None, conductivity_grid=None,
                 bounds=None, interpolation="linear"):
        if any(g is not None for g in (permittivity_grid, permeability_grid, conductivity_grid)):
            raise NotImplementedError("grid-interpolated CustomMaterial is not supported by beamz_amd")
        self.permittivity_func = permittivity_func
        self.permeability_func = permeability_func
        self.conductivity_func = conductivity_func
        self.default_permittivity = 1.0
        self.default_permeability = 1.0
        self.default_conductivity = 0.0

    def _eval(self, f, default, x, y, z):
        if f is None:
            return default
        try:
            return float(f(x, y, z))
        except TypeError:
            return float(f(x, y))

    def get_permittivity(self, x, y, z=0):
        return self._eval(self.permittivity_func, self.default_permittivity, x, y, z)

    def get_permeability(self, x, y, z=0):
        return self._eval(self.permeability_func, self.default_permeability, x, y, z)

    def get_conductivity(self, x, y, z=0):
        return self._eval(self.conductivity_func, self.default_conductivity, x, y, z)

    def get_sample(self):
        return self.default_permittivity, self.default_permeability, self.default_conductivity
