"""Float64 restatement of the reference's host point sampler.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): the checker the HIP
point sampler is compared against.

Follows the numpy path of reference ``spectral/sampling.py``:
``sample_scalar`` (:382-426), ``sample_gradient`` (:429-478) and
``sample_wind`` (:481-515): per order m, the amplitude
``sum_n c[..., n, m] Pbar_nm(sin lat)`` times ``exp(i m lon)``, real part,
doubled for m > 0.  The Legendre columns are the oracle's recurrence
(``legendre_columns`` / ``derivative_column``), which is the reference's
``legendre_bands`` / ``legendre_derivative_block`` bit for bit.
"""
from __future__ import annotations

import numpy as np

from .sht import OracleTransform, derivative_column, legendre_columns


def _points(latitude_deg, longitude_deg):
    lat, lon = np.broadcast_arrays(np.asarray(latitude_deg, dtype=np.float64),
                                   np.asarray(longitude_deg, dtype=np.float64))
    return np.deg2rad(lat.ravel()), np.deg2rad(lon.ravel()), lat.shape


def sample_scalar(o: OracleTransform, coeff, latitude_deg, longitude_deg):
    """sampling.py:382-426 (host branch)."""
    c = np.asarray(coeff, dtype=np.complex128)
    lat, lon, shape = _points(latitude_deg, longitude_deg)
    out = np.zeros((*c.shape[:-2], lat.size))
    for m, col in legendre_columns(o.truncation, np.sin(lat)):
        amplitude = np.einsum("...n,np->...p", c[..., m:, m], col)
        phase = np.exp(1j * m * lon)
        out += (1.0 if m == 0 else 2.0) * (amplitude * phase).real
    return out.reshape((*c.shape[:-2], *shape))


def sample_gradient(o: OracleTransform, coeff, latitude_deg, longitude_deg):
    """sampling.py:429-478 (host branch): (east, north) per metre."""
    c = np.asarray(coeff, dtype=np.complex128)
    lat, lon, shape = _points(latitude_deg, longitude_deg)
    x = np.sin(lat)
    zonal = np.zeros((*c.shape[:-2], lat.size))
    meridional = np.zeros_like(zonal)
    for m, col in legendre_columns(o.truncation, x):
        dcol = derivative_column(m, x, col)
        phase = np.exp(1j * m * lon)
        amp = np.einsum("...n,np->...p", c[..., m:, m], col)
        damp = np.einsum("...n,np->...p", c[..., m:, m], dcol)
        mult = 1.0 if m == 0 else 2.0
        zonal += mult * (1j * m * amp * phase).real
        meridional += mult * (damp * phase).real
    east = zonal / (o.grid.radius_m * np.cos(lat))
    north = meridional / o.grid.radius_m
    return (east.reshape((*c.shape[:-2], *shape)), north.reshape((*c.shape[:-2], *shape)))


def sample_wind(o: OracleTransform, vorticity, divergence, latitude_deg, longitude_deg):
    """sampling.py:481-515: u = -d psi/north + d chi/east, v = d psi/east + d chi/north."""
    pe, pn = sample_gradient(o, o.inverse_laplacian(vorticity), latitude_deg, longitude_deg)
    ce, cn = sample_gradient(o, o.inverse_laplacian(divergence), latitude_deg, longitude_deg)
    return -pn + ce, pe + cn
