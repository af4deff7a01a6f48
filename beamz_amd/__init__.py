"""beamz_amd: the beamz FDTD API on MI355X (gfx950) HIP kernels.

Exports the same names as ``beamz`` (beamz/__init__.py:43-107) so that
``from beamz_amd import *`` replaces ``from beamz import *`` in user scripts.
"""
from .const import EPS_0, LIGHT_SPEED, MU_0, VAC_PERMEABILITY, VAC_PERMITTIVITY, nm, um, µm, μm
from .design import CustomMaterial, Circle, CircularBend, Design, Material, PML, Polygon, Rectangle, Ring, Taper
from .devices import GaussianSource, ModeSource, Monitor, plot_signal, ramped_cosine, slab_mode_source, solve_modes
from .helpers import (calc_optimal_fdtd_params, check_fdtd_stability, code_preview, create_rich_progress,
                      display_header, display_optimization_progress, display_parameters, display_results,
                      display_simulation_status, display_status, display_time_elapsed, get_si_scale_and_label,
                      tree_view)
from .optimization.optimizers import Optimizer
from .optimization.topology import compute_overlap_gradient
from .simulation import FDTD, RegularGrid, RegularGrid3D, get_backend

__version__ = "0.1.5"

__all__ = ["LIGHT_SPEED", "VAC_PERMITTIVITY", "VAC_PERMEABILITY", "EPS_0", "MU_0", "um", "nm", "µm", "μm",
           "Material", "CustomMaterial", "Design", "Rectangle", "Circle", "Ring", "CircularBend", "Polygon", "Taper",
           "ModeSource", "GaussianSource", "Monitor", "ramped_cosine", "plot_signal", "solve_modes",
           "slab_mode_source", "RegularGrid", "FDTD", "get_backend", "Optimizer", "compute_overlap_gradient",
           "display_header", "display_status", "create_rich_progress", "display_parameters", "display_results",
           "display_simulation_status", "display_optimization_progress", "display_time_elapsed", "tree_view",
           "code_preview", "get_si_scale_and_label", "calc_optimal_fdtd_params"]
