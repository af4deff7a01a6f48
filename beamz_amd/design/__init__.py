from .core import Design
from .materials import CustomMaterial, Material
from .pml import PML
from .structures import Circle, CircularBend, Polygon, Rectangle, Ring, Taper

__all__ = ["Design", "Material", "CustomMaterial", "PML", "Polygon", "Rectangle", "Circle", "Ring",
           "CircularBend", "Taper"]
