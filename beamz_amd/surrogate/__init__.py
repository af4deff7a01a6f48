"""CNN surrogate (density map -> port transmission spectrum and its gradient) on BF16-MFMA HIP kernels."""
from .net import LEVELS, Surrogate

__all__ = ["LEVELS", "Surrogate"]
