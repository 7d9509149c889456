"""Physical constants of the transform (reference spectral/constants.py:4, 10)."""
from __future__ import annotations

EARTH_RADIUS_M = 6_371_220.0
REFERENCE_PRESSURE_PA = 100_000.0
