"""Gaussian latitude/longitude grid (reference spectral/grid.py).

Same sizing rules, exactness rules and refusals as the reference
``GaussianGrid``; the Gauss-Legendre nodes and weights come from the
engine's FP64 Newton iteration on the device (``arwen_sht_gauss_nodes``)
instead of numpy's ``leggauss`` (reference grid.py:112).  Latitudes ascend
(south to north), as ``leggauss`` orders them.
"""
from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass

import numpy as np

from . import _native
from .constants import EARTH_RADIUS_M


def _as_int(value, name: str) -> int:
    if isinstance(value, bool) or not isinstance(value, (int, np.integer)):
        raise ValueError(f"{name} must be an integer")
    return int(value)


def gauss_nodes(nlat: int) -> tuple[np.ndarray, np.ndarray]:
    """Device Newton nodes (ascending) and weights of ``nlat`` points."""
    nlat = _as_int(nlat, "nlat")
    lib = _native.load()
    mu = np.empty(nlat, dtype=np.float64)
    w = np.empty(nlat, dtype=np.float64)
    _native.check(lib.arwen_sht_gauss_nodes(nlat, _native.dptr(mu), _native.dptr(w)))
    return mu, w


@dataclass(frozen=True)
class GaussianGrid:
    truncation: int
    nlat: int
    nlon: int
    sin_lat: np.ndarray
    lat_rad: np.ndarray
    lon_rad: np.ndarray
    quadrature_weights: np.ndarray
    cos_lat: np.ndarray
    radius_m: float = EARTH_RADIUS_M

    @staticmethod
    def shape_for(truncation, *, nlat=None, nlon=None, dealias_factor: float = 1.5):
        """(nlat, nlon) for a truncation (reference grid.py:25-94): at least
        T+1 / 2T+1 and ceil(f(T+1)) / ceil(2f(T+1)) rounded up to even; an
        explicit size must meet the same minimum and nlon must be even."""
        t = _as_int(truncation, "truncation")
        if t < 1:
            raise ValueError(f"truncation must be >= 1, got {truncation}")
        if isinstance(dealias_factor, bool) or not math.isfinite(dealias_factor) \
                or dealias_factor < 1.0:
            raise ValueError("dealias_factor must be finite and >= 1")
        need_lat = max(t + 1, int(math.ceil(dealias_factor * (t + 1))))
        need_lon = max(2 * t + 1, int(math.ceil(2.0 * dealias_factor * (t + 1))))
        if nlat is None:
            nl = need_lat
        else:
            nl = _as_int(nlat, "nlat")
            if nl < need_lat:
                raise ValueError(
                    f"nlat={nl} is too small for T{t} with dealias_factor="
                    f"{dealias_factor:g}; need at least {need_lat}")
        if nlon is None:
            no = need_lon + (need_lon % 2)
        else:
            no = _as_int(nlon, "nlon")
            if no % 2:
                raise ValueError(
                    "explicit nlon must be even so positive/negative FFT "
                    "frequencies have one unambiguous layout")
            if no < need_lon:
                raise ValueError(
                    f"nlon={no} is too small for T{t} with dealias_factor="
                    f"{dealias_factor:g}; need at least {need_lon} (rounded up "
                    "to the next even value)")
        return nl, no

    @classmethod
    def create(cls, truncation, *, nlat=None, nlon=None, dealias_factor: float = 1.5,
               radius_m: float = EARTH_RADIUS_M) -> "GaussianGrid":
        nl, no = cls.shape_for(truncation, nlat=nlat, nlon=nlon, dealias_factor=dealias_factor)
        if not math.isfinite(radius_m) or radius_m <= 0:
            raise ValueError("radius_m must be finite and positive")
        mu, w = gauss_nodes(nl)
        return cls(
            truncation=int(truncation), nlat=nl, nlon=no, sin_lat=mu,
            lat_rad=np.arcsin(mu),
            lon_rad=np.arange(no, dtype=np.float64) * (2.0 * np.pi / no),
            quadrature_weights=w, cos_lat=np.sqrt(np.maximum(0.0, 1.0 - mu * mu)),
            radius_m=float(radius_m),
        )

    @classmethod
    def for_shape(cls, nlat: int, nlon: int, *, radius_m: float = EARTH_RADIUS_M):
        """A product's native grid; its truncation is nlon/2 - 1 (grid.py:128-139)."""
        return cls.create(int(nlon) // 2 - 1, nlat=nlat, nlon=nlon, dealias_factor=1.0,
                          radius_m=radius_m)

    @property
    def shape(self) -> tuple[int, int]:
        return self.nlat, self.nlon

    @property
    def zonal_content_truncation(self) -> int:
        return self.nlon // 2 - 1

    @property
    def exact_analysis_truncation(self) -> int:
        """Largest T analysing this grid's own content exactly:
        min(T_in, 2 nlat - 1 - T_in) (grid.py:157-172)."""
        t_in = self.zonal_content_truncation
        return min(t_in, 2 * self.nlat - 1 - t_in)

    def truncated_to(self, truncation: int) -> "GaussianGrid":
        """These nodes under a smaller truncation, refusing an aliasing
        projection with the condition spelt out (grid.py:174-201)."""
        t = _as_int(truncation, "truncation")
        if t < 1:
            raise ValueError(f"truncation must be >= 1, got {truncation}")
        t_in = self.zonal_content_truncation
        if t > t_in:
            raise ValueError(
                f"T{t} exceeds the zonal content of a {self.nlat}x{self.nlon} grid: "
                f"the real FFT of {self.nlon} longitudes carries orders "
                f"m <= nlon/2 - 1 = {t_in}, so order {t} does not exist there")
        if t_in + t > 2 * self.nlat - 1:
            raise ValueError(
                f"T{t} on a {self.nlat}x{self.nlon} grid is not an exact projection: "
                f"the exactness condition T_in + T <= 2*nlat - 1 with T_in = "
                f"nlon/2 - 1 = {t_in} reads {t_in} + {t} = {t_in + t} > "
                f"{2 * self.nlat - 1}, so degrees above {2 * self.nlat - 1 - t_in} "
                f"would alias; T <= {self.exact_analysis_truncation} is exact on this grid")
        return dataclasses.replace(self, truncation=t)

    @property
    def latitude_deg(self) -> np.ndarray:
        return np.rad2deg(self.lat_rad)

    @property
    def longitude_deg(self) -> np.ndarray:
        return np.rad2deg(self.lon_rad)

    def mesh(self):
        lon, lat = np.meshgrid(self.lon_rad, self.lat_rad)
        return lat, lon

    def global_mean(self, field) -> float:
        arr = _host(field)
        if arr.shape[-2:] != self.shape:
            raise ValueError(f"field has shape {arr.shape[-2:]}, expected {self.shape}")
        return float(0.5 * np.sum(arr.mean(axis=-1) * self.quadrature_weights, axis=-1))

    def global_integral(self, field) -> float:
        return 4.0 * np.pi * self.global_mean(field)

    def rms(self, field) -> float:
        arr = _host(field)
        return math.sqrt(max(0.0, self.global_mean(arr * arr)))

    def degree_length_m(self, degree: int) -> float:
        n = int(degree)
        return math.inf if n == 0 else 2.0 * np.pi * self.radius_m / n


def _host(field) -> np.ndarray:
    if hasattr(field, "detach"):
        field = field.detach().cpu().numpy()
    return np.asarray(field, dtype=np.float64)
