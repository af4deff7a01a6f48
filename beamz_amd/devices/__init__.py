from .mode import slab_mode_source, solve_modes
from .monitors import Monitor
from .signals import gaussian, gaussian_pulse, plot_signal, ramped_cosine
from .sources import GaussianSource, ModeSource

__all__ = ["GaussianSource", "ModeSource", "Monitor", "ramped_cosine", "gaussian_pulse", "gaussian", "plot_signal",
           "solve_modes", "slab_mode_source"]
