"""MI355X-native spherical-harmonic transform engine.

Drop-in for the reference's ``arwen_global.spectral`` transform layer:
``SphericalHarmonicTransform`` (analysis, synthesis, gradients),
``VorticityDivergenceOperator`` (winds <-> vorticity/divergence) and
``GaussianGrid``, computed by the HIP kernels of ``libarwen_sht.so``
(Gauss nodes and Legendre tables in FP64 on the device, parity-folded
grouped Legendre GEMMs on MFMA in FP32 or split BF16, fused radix-2/3/4
longitude FFTs).  ``TransformBatch`` runs many transforms as one grouped
launch per direction, optionally as a hipGraph.  ``FourierWaist`` is the
(*lead, nlat, T+1) buffer between the FFT and the Legendre contraction that
a latitude-band pipeline fills and drains.
"""
from __future__ import annotations

from .backend import Backend, get_backend
from .batch import TransformBatch
from .constants import EARTH_RADIUS_M, REFERENCE_PRESSURE_PA
from .ensemble import EnsembleSynthesis, ensemble_grid_fields
from .grid import GaussianGrid
from .offsets import (LNPS_REFERENCE, lnps_anomaly_coefficients, lnps_from_anomaly_coefficients,
                      theta_analysis_offsets)
from .transform import DEFAULT_BAND, LegendreTable, SphericalHarmonicTransform
from .vector import VorticityDivergenceOperator
from .waist import FourierWaist, latitude_band_edges

__all__ = [
    "Backend", "get_backend", "TransformBatch", "EARTH_RADIUS_M", "GaussianGrid",
    "DEFAULT_BAND", "SphericalHarmonicTransform", "VorticityDivergenceOperator",
    "FourierWaist", "LegendreTable", "latitude_band_edges", "REFERENCE_PRESSURE_PA",
    "LNPS_REFERENCE", "lnps_anomaly_coefficients", "lnps_from_anomaly_coefficients",
    "theta_analysis_offsets", "EnsembleSynthesis", "ensemble_grid_fields",
]
