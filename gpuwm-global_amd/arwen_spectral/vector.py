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
from .waist import FourierWaist


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

    # ---- through the Fourier waist (reference vector.py:75-165) ---------
    def vordiv_from_wind_waist(self, waist: FourierWaist):
        """:meth:`vordiv_from_wind` from the stacked ``(2, *lead, nlat, T+1)``
        waist of (u, v) that a band pipeline filled; the waist is read whole
        and kept (vector.py:75-82)."""
        return self._vordiv_from_fourier(waist)

    def _wind_fourier(self, pair, *, bands: int | None = None) -> FourierWaist:
        """The waist of a stacked ``(2, ..., nlat, nlon)`` wind pair
        (vector.py:84-111)."""
        return self.transform.fourier_waist(pair, bands=bands)

    def _pair_waists(self, fourier):
        t = self.transform
        if isinstance(fourier, FourierWaist):
            fourier.close()
            raw = fourier.values
        else:
            raw = t.backend.asarray(fourier, dtype=t.backend.complex_dtype)
        if raw.dim() < 3 or raw.shape[0] != 2:
            raise ValueError("a vector waist stacks (u, v) on the leading axis, got shape "
                             f"{tuple(raw.shape)}")
        if raw.shape[-1] < t.truncation + 1:
            raise ValueError(f"zonal spectrum has {raw.shape[-1]} orders, fewer than the "
                             f"{t.truncation + 1} the truncation keeps")
        fm = t._waist_tensor(raw[..., : t.truncation + 1])
        return fm[0], fm[1]

    def _vordiv_from_fourier(self, fourier):
        """Vorticity and divergence from the (u, v) waist, or a raw zonal
        spectrum at least T+1 orders wide, read whole in latitude; both
        projected (vector.py:113-134)."""
        t = self.transform
        uw, vw = self._pair_waists(fourier)
        zeta = t._empty_spectral(uw.shape[:-2])
        div = t._empty_spectral(uw.shape[:-2])
        _native.check(t._lib.arwen_sht_vordiv_from_waist(t._plan, _vp(uw), _vp(vw), _vp(zeta),
                                                         _vp(div), t._count(uw),
                                                         t.backend.stream()))
        return zeta, div

    def divergence_from_fourier(self, fourier):
        """The divergence alone from the (u, v) waist (vector.py:136-166): the
        zonal component against P m / cos and the meridional one against
        dP/dphi, two plane contractions per level instead of the pair's four;
        projected.  The same weights and products as the pair's divergence,
        in one grouped launch at half the columns."""
        t = self.transform
        uw, vw = self._pair_waists(fourier)
        div = t._empty_spectral(uw.shape[:-2])
        _native.check(t._lib.arwen_sht_divergence_from_waist(t._plan, _vp(uw), _vp(vw), _vp(div),
                                                             t._count(uw), t.backend.stream()))
        return div

    def flux_curl_divergence(self, eastward_flux, northward_flux):
        return self.vordiv_from_wind(eastward_flux, northward_flux)
