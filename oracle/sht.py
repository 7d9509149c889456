"""Float64 numpy restatement of the reference spherical-harmonic transform.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): the checker that the
HIP engine's results are compared against, and the timed CPU baseline.

Conventions (reference ``docs/LEVEL3_GLOBAL_SPECTRAL_DYCORE.md:42-70``):
grid fields are ``(*lead, nlat, nlon)`` real on a Gaussian grid whose
latitudes ascend (south to north, ``sin_lat`` from ``leggauss``); spectral
fields are ``(*lead, T+1, T+1)`` complex indexed ``[n, m]``, zero above the
triangle, complex-orthonormal with the Condon-Shortley phase.

Every elementwise expression in the recurrences keeps the operand order of
the reference so the float64 tables come out bit-identical to it (pinned by
the reference's recorded table digests under numpy 2.2.6).  The contraction
follows the reference numpy branch: one BLAS GEMM per order on the order's
dense (zero-prefixed) table slice.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

EARTH_RADIUS_M = 6_371_220.0  # reference spectral/constants.py:4
DEFAULT_BAND = 32  # reference spectral/legendre.py:36


# --------------------------------------------------------------------------
# Grid (reference spectral/grid.py:25-126)
# --------------------------------------------------------------------------

def grid_shape(truncation, *, nlat=None, nlon=None, dealias_factor=1.5):
    """(nlat, nlon) for a truncation; reference ``GaussianGrid.shape_for``
    (grid.py:25-94): nlat >= max(T+1, ceil(f(T+1))), nlon >= max(2T+1,
    ceil(2f(T+1))) rounded up to even; explicit sizes are checked, not
    widened."""
    if isinstance(truncation, bool) or not isinstance(truncation, (int, np.integer)):
        raise ValueError("truncation must be an integer")
    t = int(truncation)
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
        nl = int(nlat)
        if nl < need_lat:
            raise ValueError(f"nlat={nl} is too small for T{t}; need at least {need_lat}")
    if nlon is None:
        no = need_lon + (need_lon % 2)
    else:
        no = int(nlon)
        if no % 2:
            raise ValueError("explicit nlon must be even")
        if no < need_lon:
            raise ValueError(f"nlon={no} is too small for T{t}; need at least {need_lon}")
    return nl, no


@dataclass(frozen=True)
class OracleGrid:
    truncation: int
    nlat: int
    nlon: int
    sin_lat: np.ndarray
    quadrature_weights: np.ndarray
    cos_lat: np.ndarray
    lon_rad: np.ndarray
    radius_m: float = EARTH_RADIUS_M

    @property
    def shape(self):
        return self.nlat, self.nlon

    def global_mean(self, field):
        """Reference ``GaussianGrid.global_mean`` (grid.py:215-220)."""
        arr = np.asarray(field, dtype=np.float64)
        lon_mean = arr.mean(axis=-1)
        return float(0.5 * np.sum(lon_mean * self.quadrature_weights, axis=-1))


def gaussian_grid(truncation, *, nlat=None, nlon=None, dealias_factor=1.5,
                  radius_m=EARTH_RADIUS_M) -> OracleGrid:
    """Reference ``GaussianGrid.create`` (grid.py:96-126): Gauss-Legendre
    nodes ``mu`` (ascending) and weights from ``leggauss``, cos = sqrt(1-mu^2),
    longitudes 2 pi k / nlon."""
    nl, no = grid_shape(truncation, nlat=nlat, nlon=nlon, dealias_factor=dealias_factor)
    mu, w = np.polynomial.legendre.leggauss(nl)
    cos = np.sqrt(np.maximum(0.0, 1.0 - mu * mu))
    lon = np.arange(no, dtype=np.float64) * (2.0 * np.pi / no)
    return OracleGrid(int(truncation), nl, no, mu.astype(np.float64),
                      w.astype(np.float64), cos, lon, float(radius_m))


# --------------------------------------------------------------------------
# Legendre functions (reference spectral/legendre.py:76-146)
# --------------------------------------------------------------------------

def legendre_columns(truncation, x):
    """Yield ``(m, col)`` with ``col[r, j] = Pbar_{m+r}^m(x_j)`` for r = 0..T-m.

    The normalized recurrence of reference legendre.py:86-113, one order at
    a time, with the reference's operand order in every expression:
    diagonal  P_mm = (-sqrt((2m+1)/(2m)) * root) * P_{m-1,m-1}
    first     P_{m+1,m} = (sqrt(2m+3) * x) * P_mm
    interior  P_nm = a * (x * P_{n-1,m} - b * P_{n-2,m}).
    """
    t = int(truncation)
    x = np.asarray(x, dtype=np.float64)
    root = np.sqrt(np.maximum(0.0, 1.0 - x * x))
    diag = np.full(x.shape, 1.0 / math.sqrt(4.0 * math.pi))
    for m in range(t + 1):
        if m > 0:
            diag = -math.sqrt((2 * m + 1) / (2.0 * m)) * root * diag
        col = np.zeros((t + 1 - m, x.size), dtype=np.float64)
        col[0] = diag
        if m < t:
            col[1] = math.sqrt(2 * m + 3) * x * diag
        for n in range(m + 2, t + 1):
            a = math.sqrt((4.0 * n * n - 1.0) / (n * n - m * m))
            b = math.sqrt(((n - 1.0) * (n - 1.0) - m * m)
                          / (4.0 * (n - 1.0) * (n - 1.0) - 1.0))
            r = n - m
            col[r] = a * (x * col[r - 1] - b * col[r - 2])
        yield m, col


def derivative_column(m, x, col):
    """``dPbar_n^m/dphi`` for one order's column (reference legendre.py:120-146).

    ``(x^2-1) dP/dx = n x P_nm - c_nm P_{n-1,m}`` with
    ``c_nm = sqrt((n^2-m^2)(2n+1)/(2n-1))``; dP/dphi = cos * that / (x^2-1).
    """
    x = np.asarray(x, dtype=np.float64)
    if x.ndim != 1 or np.any(np.abs(x) >= 1.0):
        raise ValueError("derivatives need latitudes strictly inside (-1, 1)")
    denom = x * x - 1.0
    cos = np.sqrt(np.maximum(0.0, 1.0 - x * x))
    out = np.zeros_like(col)
    if m >= 1:
        out[0] = cos * (m * x * col[0]) / denom
    if col.shape[0] > 1:
        n = np.arange(m + 1, m + col.shape[0], dtype=np.int64)
        c = np.sqrt((n * n - m * m) * (2.0 * n + 1.0) / (2.0 * n - 1.0))
        num = n[:, None] * x[None, :] * col[1:] - c[:, None] * col[:-1]
        out[1:] = cos[None, :] * num / denom[None, :]
    return out


def analysis_rows(col, weights):
    """Gram-solved analysis rows of one order (reference transform.py:97-119).

    rows = (P * w) * 2 pi; A = solve(rows @ P^T, rows).  Returns
    ``(nlat, T+1-m)`` (the reference's 'jn' layout for one order).
    """
    rows = col * weights[None, :] * (2.0 * math.pi)
    gram = rows @ col.T
    return np.linalg.solve(gram, rows).T


def dense_tables(truncation, x, weights=None, *, derivative=True):
    """The reference dense-reference tables (``_basis_mnj``, ``_dbasis_mnj``
    (m, n, j) and ``_analysis_mjn`` (m, j, n)), float64.  Test use at small T.
    """
    t = int(truncation)
    x = np.asarray(x, dtype=np.float64)
    basis = np.zeros((t + 1, t + 1, x.size))
    dbasis = np.zeros_like(basis) if derivative else None
    ana = np.zeros((t + 1, x.size, t + 1)) if weights is not None else None
    for m, col in legendre_columns(t, x):
        basis[m, m:] = col
        if derivative:
            dbasis[m, m:] = derivative_column(m, x, col)
        if ana is not None:
            ana[m, :, m:] = analysis_rows(col, weights)
    return basis, dbasis, ana


# --------------------------------------------------------------------------
# Transform (reference spectral/transform.py, numpy branches)
# --------------------------------------------------------------------------

class OracleTransform:
    """Float64 spherical-harmonic transform, reference numpy semantics.

    Tables are packed per order (``[m]`` -> ``(T+1-m, nlat)``) and expanded to
    the dense zero-prefixed slice per band for the GEMM, as the reference's
    ``BandedLegendreTable.expanded`` does (transform.py:122-212), so each
    order's BLAS call sees the reference's operand shapes.
    """

    def __init__(self, grid: OracleGrid, band: int = DEFAULT_BAND):
        self.grid = grid
        self.band = int(band)
        t = grid.truncation
        self.truncation = t
        self._basis = []
        self._analysis = []
        for m, col in legendre_columns(t, grid.sin_lat):
            self._basis.append(col)
            self._analysis.append(analysis_rows(col, grid.quadrature_weights))
        self._dbasis = None
        n = np.arange(t + 1, dtype=np.float64)
        self._eigen = -(n * (n + 1.0)) / (grid.radius_m * grid.radius_m)
        self._zonal = 1j * np.arange(t + 1, dtype=np.float64)
        self._tri = np.tri(t + 1, dtype=bool)
        scale = 2.0 * math.pi / grid.radius_m
        m = np.arange(t + 1, dtype=np.float64)
        # reference vector.py:16-34
        self._rotation_weight = ((1j * m)[None, :]
                                 * (grid.quadrature_weights / grid.cos_lat * scale)[:, None])
        self._shear_weight = (grid.quadrature_weights * scale)[:, None]

    @classmethod
    def create(cls, truncation, *, nlat=None, nlon=None, dealias_factor=1.5,
               radius_m=EARTH_RADIUS_M, band=DEFAULT_BAND):
        return cls(gaussian_grid(truncation, nlat=nlat, nlon=nlon,
                                 dealias_factor=dealias_factor, radius_m=radius_m),
                   band=band)

    # ---- shapes ---------------------------------------------------------
    @property
    def spectral_shape(self):
        return self.truncation + 1, self.truncation + 1

    def _check_grid(self, f):
        if tuple(f.shape[-2:]) != self.grid.shape:
            raise ValueError(f"grid field has trailing shape {f.shape[-2:]}, "
                             f"expected {self.grid.shape}")

    def _check_spectral(self, c):
        if tuple(c.shape[-2:]) != self.spectral_shape:
            raise ValueError(f"spectral field has trailing shape {c.shape[-2:]}, "
                             f"expected {self.spectral_shape}")

    @property
    def dbasis(self):
        if self._dbasis is None:
            x = self.grid.sin_lat
            self._dbasis = [derivative_column(m, x, col) for m, col in enumerate(self._basis)]
        return self._dbasis

    # ---- the contraction (transform.py:1178-1274, numpy branch) ---------
    def _bands(self):
        t1 = self.truncation + 1
        return [(m0, min(m0 + self.band, t1)) for m0 in range(0, t1, self.band)]

    def _expand(self, packed, m0, m1, layout):
        """Dense ``(m1-m0, K, N)`` slice: 'nj' -> (T+1, nlat), 'jn' -> (nlat, T+1)."""
        t1 = self.truncation + 1
        nlat = self.grid.nlat
        if layout == "nj":
            dense = np.zeros((m1 - m0, t1, nlat))
            for i, m in enumerate(range(m0, m1)):
                dense[i, m:, :] = packed[m]
        else:
            dense = np.zeros((m1 - m0, nlat, t1))
            for i, m in enumerate(range(m0, m1)):
                dense[i, :, m:] = packed[m]
        return dense

    def _contract(self, values, packed, layout, *, transposed=False):
        """out[..., n, m] = sum_k values[..., k, m] table[m, k, n]."""
        lead = values.shape[:-2]
        k = values.shape[-2]
        mc = values.shape[-1]
        stacked = np.moveaxis(values, -1, 0).reshape(mc, -1, k)
        re_in = np.ascontiguousarray(stacked.real)
        im_in = np.ascontiguousarray(stacked.imag)
        ncount = self.truncation + 1 if (layout == "jn" or transposed) else self.grid.nlat
        re = np.empty((mc, stacked.shape[1], ncount))
        im = np.empty_like(re)
        for m0, m1 in self._bands():
            dense = self._expand(packed, m0, m1, layout)
            if transposed:
                dense = dense.transpose(0, 2, 1)
            re[m0:m1] = re_in[m0:m1] @ dense
            im[m0:m1] = im_in[m0:m1] @ dense
        out = np.empty(re.shape, dtype=np.complex128)
        out.real = re
        out.imag = im
        return np.moveaxis(out.reshape(mc, *lead, ncount), 0, -1)

    # ---- grid <-> waist (transform.py:1294-1404, 1654-1702) -------------
    def fourier_waist(self, field):
        f = np.asarray(field, dtype=np.float64)
        self._check_grid(f)
        return np.divide(np.fft.rfft(f, axis=-1)[..., : self.truncation + 1],
                         self.grid.nlon)

    def waist_to_grid(self, values):
        spectrum = np.zeros((*values.shape[:-2], values.shape[-2], self.grid.nlon // 2 + 1),
                            dtype=np.complex128)
        np.multiply(self.grid.nlon, values, out=spectrum[..., : self.truncation + 1])
        return np.fft.irfft(spectrum, n=self.grid.nlon, axis=-1).astype(np.float64)

    # ---- public operations ---------------------------------------------
    def contract_waist(self, values):
        """The tail of an analysis (transform.py:1435-1472): contract a
        (*lead, nlat, T+1) waist with the analysis table, realify m = 0."""
        out = self._contract(np.asarray(values, dtype=np.complex128), self._analysis, "jn")
        out[..., :, 0] = out[..., :, 0].real
        return out

    def forward(self, field):
        """Analysis (transform.py:1474-1490): waist, contract, realify m=0."""
        return self.contract_waist(self.fourier_waist(field))

    def contract_to_waist(self, coeff, table="basis"):
        c = np.asarray(coeff, dtype=np.complex128)
        self._check_spectral(c)
        packed = self._basis if table == "basis" else self.dbasis
        return self._contract(c, packed, "nj")

    def inverse(self, coeff):
        """Synthesis (transform.py:1629-1746)."""
        return self.waist_to_grid(self.contract_to_waist(coeff))

    def inverse_meridional_derivative(self, coeff):
        return self.waist_to_grid(self.contract_to_waist(coeff, "derivative"))

    def inverse_zonal_derivative(self, coeff):
        c = np.asarray(coeff, dtype=np.complex128)
        self._check_spectral(c)
        return self.inverse(c * self._zonal)

    def gradient(self, coeff):
        """(east, north) per metre (transform.py:1762-1768)."""
        east = self.inverse_zonal_derivative(coeff) / (
            self.grid.radius_m * self.grid.cos_lat[:, None])
        north = self.inverse_meridional_derivative(coeff) / self.grid.radius_m
        return east, north

    def gradient_waists(self, coeff):
        """(contract_to_waist(c i m), contract_to_waist(c, dP/dphi))
        (transform.py:1770-1787)."""
        c = np.asarray(coeff, dtype=np.complex128)
        self._check_spectral(c)
        return self.contract_to_waist(c * self._zonal), self.contract_to_waist(c, "derivative")

    def gradient_band(self, waists, r0, r1):
        """Rows r0..r1 of gradient() drained from its waists (transform.py:1789-1796)."""
        ew, nw = waists
        east = self.waist_to_grid(ew[..., r0:r1, :]) / (
            self.grid.radius_m * self.grid.cos_lat[r0:r1, None])
        north = self.waist_to_grid(nw[..., r0:r1, :]) / self.grid.radius_m
        return east, north

    def project(self, coeff):
        """transform.py:1124-1144 (numpy branch)."""
        arr = np.asarray(coeff, dtype=np.complex128).copy()
        self._check_spectral(arr)
        arr *= self._tri
        arr[..., :, 0] = arr[..., :, 0].real
        return arr

    def laplacian(self, coeff):
        c = np.asarray(coeff, dtype=np.complex128)
        self._check_spectral(c)
        return c * self._eigen[:, None]

    def inverse_laplacian(self, coeff):
        c = np.asarray(coeff, dtype=np.complex128)
        self._check_spectral(c)
        out = np.zeros_like(c)
        out[..., 1:, :] = c[..., 1:, :] / self._eigen[1:, None]
        return self.project(out)

    def constant_coeff(self, value):
        out = np.zeros(self.spectral_shape, dtype=np.complex128)
        out[0, 0] = float(value) * math.sqrt(4.0 * math.pi)
        return out

    def spectral_mean_square(self, coeff):
        c = np.asarray(coeff)
        w = np.ones(self.truncation + 1)
        w[1:] = 2.0
        total = np.sum(np.abs(c) ** 2 * w[None, :], axis=(-2, -1))
        return float(np.asarray(total).mean() / (4.0 * np.pi))

    def power_by_degree(self, coeff):
        c = np.asarray(coeff)
        w = np.ones(self.truncation + 1)
        w[1:] = 2.0
        return np.sum(np.abs(c) ** 2 * w[None, :], axis=-1)

    def transform_check(self, *, seed=0, max_degree=None):
        """transform.py:1837-1863."""
        rng = np.random.default_rng(seed)
        t = self.truncation if max_degree is None else min(self.truncation, int(max_degree))
        coeff = np.zeros(self.spectral_shape, dtype=np.complex128)
        for n in range(t + 1):
            for m in range(n + 1):
                coeff[n, m] = rng.normal() + (0.0j if m == 0 else 1j * rng.normal())
        grid = self.inverse(coeff)
        back = self.forward(grid)
        scale = max(1.0, float(np.max(np.abs(coeff))))
        error = float(np.max(np.abs(back - coeff)) / scale)
        spatial = self.grid.global_mean(grid ** 2)
        spectral = float(np.sum(np.abs(coeff) ** 2 * np.where(
            np.arange(self.truncation + 1)[None, :] == 0, 1.0, 2.0)) / (4.0 * np.pi))
        parseval = abs(spatial - spectral) / max(1.0, abs(spectral))
        return {"truncation": self.truncation, "nlat": self.grid.nlat,
                "nlon": self.grid.nlon, "roundtrip_relative_linf": error,
                "parseval_relative_error": float(parseval)}

    # ---- vector operator (reference spectral/vector.py) -----------------
    def wind_from_vordiv(self, vorticity, divergence):
        """vector.py:36-52."""
        psi = self.inverse_laplacian(vorticity)
        chi = self.inverse_laplacian(divergence)
        east, north = self.gradient(np.stack([psi, chi]))
        return -north[0] + east[1], east[0] + north[1]

    def vordiv_from_wind(self, u, v):
        """vector.py:54-73, 113-134: integration by parts."""
        u = np.asarray(u, dtype=np.float64)
        v = np.asarray(v, dtype=np.float64)
        self._check_grid(u)
        if u.shape != v.shape:
            raise ValueError(f"u/v shape mismatch: {u.shape} != {v.shape}")
        return self.vordiv_from_waist(self.fourier_waist(np.stack([u, v])))

    def vordiv_from_waist(self, fourier):
        """vector.py:113-134: the contractions of a stacked (2, ..., nlat, >= T+1)
        (u, v) waist or raw zonal spectrum."""
        fm = np.asarray(fourier, dtype=np.complex128)[..., : self.truncation + 1]
        rot = self._contract(fm * self._rotation_weight, self._basis, "nj", transposed=True)
        sh = self._contract(fm * self._shear_weight, self.dbasis, "nj", transposed=True)
        div = rot[0] - sh[1]
        zeta = rot[1] + sh[0]
        return self.project(zeta), self.project(div)


# --------------------------------------------------------------------------
# Synthetic inputs: seeded red-spectrum coefficients
# --------------------------------------------------------------------------

def red_spectrum(truncation, lead=(), *, seed=0, scale=1.0, mean=None, slope=1.5):
    """c_nm = scale * n^-slope * (N(0,1) + i N(0,1)) / sqrt(2) for 1 <= n <= T,
    0 <= m <= n; m = 0 imaginary parts zeroed; zero above the triangle;
    (0, 0) = sqrt(4 pi) * mean when a mean is given, else 0.
    Generator: numpy PCG64(seed)."""
    t1 = int(truncation) + 1
    rng = np.random.Generator(np.random.PCG64(seed))
    lead = tuple(int(v) for v in lead)
    re = rng.standard_normal((*lead, t1, t1))
    im = rng.standard_normal((*lead, t1, t1))
    n = np.arange(t1, dtype=np.float64)
    amp = np.zeros(t1)
    amp[1:] = n[1:] ** (-slope)
    c = (re + 1j * im) * (scale * amp / math.sqrt(2.0))[:, None]
    c *= np.tri(t1, dtype=bool)
    c[..., :, 0] = c[..., :, 0].real
    if mean is not None:
        c[..., 0, 0] = math.sqrt(4.0 * math.pi) * float(mean)
    return c
