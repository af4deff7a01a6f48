"""Physical constants and units, with the exact values of beamz/const.py:1-10.

The permittivity and permeability of vacuum are the reference's rounded values
(not CODATA); they enter every update coefficient, so they are part of parity.
"""
LIGHT_SPEED = 299792458.0
VAC_PERMITTIVITY = 8.85e-12
VAC_PERMEABILITY = 1.256e-6
EPS_0 = 8.85e-12
MU_0 = 1.256e-6

um = 1e-6
µm = 1e-6
μm = 1e-6
nm = 1e-9
