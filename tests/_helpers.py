"""Shared helpers for the GPU parity tests (no test functions here)."""
from __future__ import annotations

import numpy as np

from oracle.sht import OracleTransform, gaussian_grid


def hip_transform(oracle: OracleTransform, precision: str = "float32", device_nodes=False):
    """The engine's transform on the oracle's grid (leggauss nodes) or on the
    same shape with the engine's own device Newton nodes."""
    from arwen_spectral import GaussianGrid, SphericalHarmonicTransform, get_backend

    g = oracle.grid
    if device_nodes:
        grid = GaussianGrid.create(g.truncation, nlat=g.nlat, nlon=g.nlon, dealias_factor=1.0,
                                   radius_m=g.radius_m)
    else:
        grid = GaussianGrid(truncation=g.truncation, nlat=g.nlat, nlon=g.nlon,
                            sin_lat=g.sin_lat, lat_rad=np.arcsin(g.sin_lat), lon_rad=g.lon_rad,
                            quadrature_weights=g.quadrature_weights, cos_lat=g.cos_lat,
                            radius_m=g.radius_m)
    return SphericalHarmonicTransform(grid, get_backend("hip", precision))


def rel_err(have, want) -> float:
    have = np.asarray(have.detach().cpu().numpy() if hasattr(have, "detach") else have)
    want = np.asarray(want)
    scale = float(np.max(np.abs(want)))
    return float(np.max(np.abs(have.astype(want.dtype) - want)) / (scale if scale > 0 else 1.0))
