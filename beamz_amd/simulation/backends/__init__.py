"""Backend registry (beamz/simulation/backends/__init__.py:8-36).

``"hip"`` is this package's backend. The reference's names ``"numpy"``,
``"torch"`` and ``"jax"`` select the same HIP backend, so unmodified scripts
(whose default is "numpy") run on the GPU; an unknown name raises ValueError
like the reference. There is no CPU fallback: without a GPU, construction
raises RuntimeError.
"""
import logging

logger = logging.getLogger(__name__)

_ALIASES = ("hip", "numpy", "torch", "jax", "gpu", "rocm")


def get_backend(name="hip", **kwargs):
    key = name.lower()
    if key not in _ALIASES:
        raise ValueError(f"Unknown backend: {name}")
    if key != "hip":
        logger.info("backend %r runs on the beamz_amd HIP backend", name)
    from .hip_backend import HipBackend
    return HipBackend(**kwargs)


__all__ = ["get_backend"]
