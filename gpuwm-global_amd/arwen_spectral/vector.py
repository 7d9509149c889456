"""Vector spherical-harmonic transforms through vorticity/divergence.

Drop-in for the reference ``VorticityDivergenceOperator``
(src/arwen_global/spectral/vector.py).  Both directions are single engine
calls: the inverse Laplacian, the two Legendre tables and the u/v assembly
are fused into one grouped GEMM and its inverse FFTs (wind_from_vordiv), and
the quadrature/metric weights, the integration by parts and the projection
into the forward FFT, one grouped GEMM and the unpack (vordiv_from_wind).
"""
from __future__ import annotations

from dataclasses import dataclass

from . import _native
from .transform import SphericalHarmonicTransform, _vp


@dataclass
class VorticityDivergenceOperator:
    transform: SphericalHarmonicTransform

    def wind_from_vordiv(self, vorticity, divergence):
        """u, v from vorticity/divergence via psi, chi (reference vector.py:36-52):
        u = -dpsi/dphi / a + dchi/dlambda / (a cos), v = dpsi/dlambda / (a cos) + dchi/dphi / a."""
        t = self.transform
        z = t._spectral_tensor(vorticity)
        d = t._spectral_tensor(divergence)
        if z.shape != d.shape:
            raise ValueError(f"vorticity/divergence shape mismatch: {tuple(z.shape)} != {tuple(d.shape)}")
        u = t._empty_grid(z.shape[:-2])
        v = t._empty_grid(z.shape[:-2])
        _native.check(t._lib.arwen_sht_uv_from_vordiv(t._plan, _vp(z), _vp(d), _vp(u), _vp(v),
                                                      t._count(z), t.backend.stream()))
        return u, v

    def vordiv_from_wind(self, u, v):
        """Vorticity and divergence of a grid wind by integration by parts
        (reference vector.py:54-73, 113-134); both outputs projected."""
        t = self.transform
        uu = t._grid_tensor(u)
        vv = t._grid_tensor(v)
        if uu.shape != vv.shape:
            raise ValueError(f"u/v shape mismatch: {tuple(uu.shape)} != {tuple(vv.shape)}")
        zeta = t._empty_spectral(uu.shape[:-2])
        div = t._empty_spectral(uu.shape[:-2])
        _native.check(t._lib.arwen_sht_vordiv_from_uv(t._plan, _vp(uu), _vp(vv), _vp(zeta),
                                                      _vp(div), t._count(uu), t.backend.stream()))
        return zeta, div

    def flux_curl_divergence(self, eastward_flux, northward_flux):
        return self.vordiv_from_wind(eastward_flux, northward_flux)
