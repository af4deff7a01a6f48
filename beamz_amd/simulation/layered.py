"""Layer-stack 3D meshes built from the bit-exact 2D rasterizer.

The reference's 3D rasterizer touches at most three z-layers per structure
and ignores depth (beamz/simulation/meshing.py:682-694), so it cannot build a
220 nm silicon slab; it also has no z-PML (meshing.py:766-779). For layered
photonic stacks this module builds the 3D maps by a stated rule instead:

* each layer is (z_min, z_max, 2D design); the 2D design is rasterized with
  ``RegularGrid`` (same painter's algorithm, super-sampling and x/y PML as the
  reference, bit-exact), once per layer;
* a cell whose z-extent straddles layer boundaries gets the z-fill-weighted
  mean of the layers' permittivity (conductivity likewise);
* x/y PML conductivity comes from each layer's own 2D raster (so it uses that
  layer's permittivity, as pml.py:31-34 does); an optional z-PML of
  ``z_pml_cells`` cells adds the same profile (pml.py:21-27) along z, with
  sigma_max = sigma_factor * 1.2 / (eta dx) from the cell's permittivity.

The FDTD driver accepts the result like any provided mesh
(beamz/simulation/fdtd.py:28-34).
"""
from __future__ import annotations

import numpy as np

from ..const import EPS_0, MU_0
from .meshing import RegularGrid, RegularGrid3D


def _pml_profile(d, order):
    d = np.clip(d, 0.0, 1.0)
    s = np.where(d < 0.05, 0.01 * (d / 0.05) ** 2, ((d - 0.05) / 0.95) ** order)
    return s / (1.0 + 0.1 * d ** 2)


def layered_grid_3d(design3d, layers, resolution, resolution_z=None, z_pml_cells=0, z_pml_order=3.0,
                    z_pml_sigma_factor=1.0):
    """Build a RegularGrid3D for ``design3d`` from ``layers`` = [(z_min, z_max, design2d), ...]."""
    rz = resolution if resolution_z is None else resolution_z
    nz = int(design3d.depth / rz)
    ny, nx = int(design3d.height / resolution), int(design3d.width / resolution)
    eps = np.zeros((nz, ny, nx))
    sig = np.zeros((nz, ny, nx))
    cover = np.zeros(nz)
    for z0, z1, d2 in layers:
        g = RegularGrid(d2, resolution)
        if g.permittivity.shape != (ny, nx):
            raise ValueError("layer designs must match the 3D design's x/y extent")
        for k in range(nz):
            lo, hi = k * rz, (k + 1) * rz
            f = max(0.0, min(hi, z1) - max(lo, z0)) / rz
            if f > 0:
                eps[k] += f * g.permittivity
                sig[k] += f * g.conductivity
                cover[k] += f
    if np.any(np.abs(cover - 1.0) > 1e-9):
        raise ValueError("layers must cover the design depth exactly once")
    if z_pml_cells > 0:
        zc = (np.arange(nz) + 0.5) * rz
        L = z_pml_cells * rz
        depth = np.maximum((L - zc) / L, (zc - (nz * rz - L)) / L)
        prof = np.where(depth > 0, _pml_profile(np.maximum(depth, 0.0), z_pml_order), 0.0)
        eta = np.sqrt(MU_0 / (EPS_0 * eps))
        sig += (z_pml_sigma_factor * 1.2 / (eta * resolution)) * prof[:, None, None]
    return RegularGrid3D.from_arrays(design3d, eps, sig, resolution, rz)


def pml_sigma_factor(cells, reflection=1e-6, order=3.0):
    """sigma_factor giving a cubic-graded layer of ``cells`` cells the target normal reflection.

    The reference profile peaks at sigma_factor * 1.2 / (eta dx); a polynomial grading of order m
    reaches R = exp(-2 (m+1)^-1 sigma_max eta d) with d = cells*dx, so
    sigma_factor = (m+1) ln(1/R) / (2 * 1.2 * cells).
    """
    return (order + 1.0) * np.log(1.0 / reflection) / (2.0 * 1.2 * cells)
