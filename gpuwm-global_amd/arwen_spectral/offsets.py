"""Offsets the dycore's float32 step takes fields about before a transform.

ln ps about the fixed ln(p_ref) (reference lnps_reference.py:1-49) and the
theta analysis about the reference column rounded to whole kelvin
(reference semilag/step.py:128-163).  The helpers build the per-field
offsets; ``SphericalHarmonicTransform.forward_about`` / ``inverse_about`` and
the batch options ``grid_offset`` / ``coeff_offset`` apply them inside the
fold and pack/unpack kernels, so no extra pass over the grid or the
coefficients is made.
"""
from __future__ import annotations

import math

import numpy as np

from .constants import REFERENCE_PRESSURE_PA

__all__ = [
    "LNPS_REFERENCE",
    "lnps_anomaly_coefficients",
    "lnps_from_anomaly_coefficients",
    "lnps_offset",
    "theta_analysis_offsets",
]

#: ln of the reference surface pressure (a fixed constant, so a restart takes
#: off and puts back the same bytes as a continuous run).
LNPS_REFERENCE = math.log(REFERENCE_PRESSURE_PA)


def lnps_offset(sign: float = 1.0) -> float:
    """``sign * ln(p_ref) * sqrt(4 pi)``: the (0, 0) coefficient that
    ``add_grid_constant(c, sign * LNPS_REFERENCE)`` adds."""
    return float(sign) * LNPS_REFERENCE * math.sqrt(4.0 * math.pi)


def lnps_anomaly_coefficients(transform, coefficients):
    """``coefficients`` with ln(p_ref) taken off (0, 0) (reference
    lnps_reference.py:32-44); exact in float32 by Sterbenz's lemma for any
    global-mean surface pressure above a few hPa."""
    return transform.add_grid_constant(coefficients, -LNPS_REFERENCE)


def lnps_from_anomaly_coefficients(transform, coefficients):
    """The inverse of :func:`lnps_anomaly_coefficients` (lnps_reference.py:47-49)."""
    return transform.add_grid_constant(coefficients, LNPS_REFERENCE)


def theta_analysis_offsets(transform, theta_reference_column):
    """``(grid, coefficient)`` per level: the reference column rounded to
    whole kelvin (float32, shape (nlev,)) and the same numbers times
    sqrt(4 pi) cast once to complex64 (reference semilag/step.py:128-157).
    ``forward_about(theta, grid, coefficient)`` is the reference's
    ``forward(theta - grid[:, None, None])`` with ``coefficient`` added at
    (0, 0) (semilag/step.py:160-163, 610-621): a constant plane then leaks
    only its departure from the offset into the m = 0 coefficients."""
    backend = transform.backend
    torch = backend.xp
    whole = np.round(np.asarray(theta_reference_column, dtype=np.float64).reshape(-1))
    grid = torch.as_tensor(whole, dtype=backend.float_dtype, device=backend.device())
    coefficient = torch.as_tensor(whole * math.sqrt(4.0 * math.pi),
                                  dtype=backend.complex_dtype, device=backend.device())
    return grid, coefficient
