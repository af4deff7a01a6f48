from .backends import get_backend
from .fdtd import FDTD
from .meshing import BaseMeshGrid, RegularGrid, RegularGrid3D

__all__ = ["RegularGrid", "RegularGrid3D", "BaseMeshGrid", "FDTD", "get_backend"]
