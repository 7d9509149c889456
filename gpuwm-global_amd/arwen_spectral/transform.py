"""Triangular spherical-harmonic analysis/synthesis on a Gaussian grid, on MI355X.

Drop-in for the reference's ``SphericalHarmonicTransform``
(src/arwen_global/spectral/transform.py:632-1863): same constructor, same
layouts -- grid ``(*lead, nlat, nlon)`` real, spectral ``(*lead, T+1, T+1)``
complex indexed ``[n, m]`` -- same method names, identity rules and error
types.  Arrays are torch tensors on the current ROCm device; numpy inputs
are copied there.  Every transform is one call into ``libarwen_sht.so``.
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import math
from dataclasses import dataclass, field

import numpy as np

from . import _native
from .backend import Backend, get_backend
from .constants import EARTH_RADIUS_M
from .grid import GaussianGrid
from .waist import FourierWaist, latitude_band_edges

DEFAULT_BAND = 32  # reference legendre.py:36; accepted for API compatibility


def _vp(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr() if t is not None else None)


@dataclass(frozen=True)
class LegendreTable:
    """Names one of a plan's device-resident Legendre tables: ``analysis``
    (P w in the analysis layout), ``basis`` (P) or ``derivative`` (dP/dphi).
    The reference hands its contractions table objects
    (transform.py:122-216, 902-934); here the tables live inside the plan
    and these names select them."""
    name: str


@dataclass
class SphericalHarmonicTransform:
    grid: GaussianGrid
    backend: Backend
    tensor_core_contractions: bool = False
    legendre_band: int = DEFAULT_BAND
    streaming: bool = False
    latitude_bands: int = 1
    order_partition: tuple = ()
    #: the vector operators' zonal orders m < fp64_orders run in FP64 end to
    #: end (arwen_sht_plan_desc.fp64_orders): None = the default (0 for
    #: float32, 4 ceil((T+1)/192) for the split-BF16 modes), 0 = off
    fp64_orders: int | None = None
    _plan: ctypes.c_void_p = field(default=None, init=False, repr=False)

    def __post_init__(self) -> None:
        if self.tensor_core_contractions:
            raise ValueError(
                "tensor_core_contractions=True selects TF32 on CUDA; on this engine "
                "choose precision='bf16x6' or 'float32' instead")
        if int(self.legendre_band) < 1:
            raise ValueError(f"legendre_band must be >= 1, got {self.legendre_band}")
        if int(self.latitude_bands) < 1:
            raise ValueError(f"latitude_bands must be >= 1, got {self.latitude_bands}")
        if self.streaming:
            raise ValueError(
                "streaming=True is not built: on 288 GB of HBM the T767 tables "
                "are resident (about 4 GB)")
        if tuple(self.order_partition or ()):
            raise ValueError("order_partition is a multi-device axis; this engine runs on one GPU")
        lib = _native.load()
        g = self.grid
        mu = np.ascontiguousarray(g.sin_lat, dtype=np.float64)
        w = np.ascontiguousarray(g.quadrature_weights, dtype=np.float64)
        fp64 = -1 if self.fp64_orders is None else int(self.fp64_orders)
        desc = _native.PlanDesc(int(g.truncation), int(g.nlat), int(g.nlon), float(g.radius_m),
                                self.backend.precision_code, _native.dptr(mu), _native.dptr(w),
                                fp64)
        plan = ctypes.c_void_p()
        _native.check(lib.arwen_sht_plan_create(ctypes.byref(desc), ctypes.byref(plan)))
        self._plan = plan
        self._lib = lib
        self._fp64_orders = int(lib.arwen_sht_plan_fp64_orders(plan))
        torch = self.backend.xp
        n = np.arange(g.truncation + 1, dtype=np.float64)
        self._eigen = -(n * (n + 1.0)) / (g.radius_m * g.radius_m)
        self._analysis = LegendreTable("analysis")
        self._basis = LegendreTable("basis")
        self._derivative_basis = LegendreTable("derivative")
        self._coslat = torch.tensor(np.asarray(g.cos_lat, dtype=np.float64)[:, None],
                                    dtype=self.backend.float_dtype, device=self.backend.device())

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan:
            try:
                self._lib.arwen_sht_plan_destroy(plan)
            except Exception:
                pass
            self._plan = None

    # ---- construction (reference transform.py:953-1023) -----------------
    @classmethod
    def create(cls, truncation: int, *, nlat=None, nlon=None, dealias_factor: float = 1.5,
               radius_m=None, grid: GaussianGrid | None = None, backend: str = "hip",
               precision: str = "float32", tensor_core_contractions: bool = False,
               legendre_band: int = DEFAULT_BAND, streaming: bool = False,
               latitude_bands: int = 1, order_partition: tuple = (),
               fp64_orders: int | None = None):
        if grid is not None:
            if nlat is not None or nlon is not None or radius_m is not None \
                    or dealias_factor != 1.5:
                raise ValueError(
                    "grid= carries its own shape and radius: nlat, nlon, "
                    "dealias_factor and radius_m cannot be combined with it")
            grid = grid.truncated_to(truncation)
        else:
            kwargs = {} if radius_m is None else {"radius_m": radius_m}
            grid = GaussianGrid.create(truncation, nlat=nlat, nlon=nlon,
                                       dealias_factor=dealias_factor, **kwargs)
        return cls(grid, get_backend(backend, precision),
                   tensor_core_contractions=tensor_core_contractions,
                   legendre_band=legendre_band, streaming=streaming,
                   latitude_bands=latitude_bands, order_partition=tuple(order_partition or ()),
                   fp64_orders=fp64_orders)

    # ---- shapes and identity (transform.py:1044-1104) -------------------
    @property
    def truncation(self) -> int:
        return self.grid.truncation

    @property
    def spectral_shape(self) -> tuple[int, int]:
        n = self.truncation + 1
        return n, n

    @property
    def plan_handle(self) -> ctypes.c_void_p:
        return self._plan

    @property
    def geometry_identity(self) -> dict:
        return {
            "truncation": int(self.truncation),
            "nlat": int(self.grid.nlat),
            "nlon": int(self.grid.nlon),
            "radius_m": float(self.grid.radius_m),
            "normalization": "complex-orthonormal-condon-shortley",
            "latitude_grid": "gauss-legendre-in-sin-latitude",
        }

    @property
    def geometry_hash(self) -> str:
        return hashlib.sha256(json.dumps(self.geometry_identity, sort_keys=True,
                                         separators=(",", ":")).encode()).hexdigest()

    @property
    def identity(self) -> dict:
        """Geometry plus arithmetic: a precision mode other than float32 joins
        the identity, as the reference joins TF32 (transform.py:1074-1098)."""
        payload = {**self.geometry_identity, "backend": self.backend.name,
                   "float_dtype": "float32"}
        if self.backend.precision != "float32":
            payload["legendre_precision"] = self.backend.precision
        if self._fp64_orders:
            payload["vector_fp64_orders"] = self._fp64_orders
        if int(self.legendre_band) != DEFAULT_BAND:
            payload["legendre_band"] = int(self.legendre_band)
        return payload

    @property
    def identity_hash(self) -> str:
        return hashlib.sha256(json.dumps(self.identity, sort_keys=True,
                                         separators=(",", ":")).encode()).hexdigest()

    @property
    def vector_fp64_orders(self) -> int:
        """The zonal orders below which the vector operators run in FP64 (0: none),
        as the plan resolved ``fp64_orders``."""
        return self._fp64_orders

    @property
    def legendre_table_nbytes(self) -> dict:
        total = int(self._lib.arwen_sht_plan_table_bytes(self._plan))
        jhp = -(-((self.grid.nlat + 1) // 2) // 32) * 32
        low = 4 * self._fp64_orders * (self.truncation + 1) * jhp * 8
        sixth = (total - low) // 6  # P, dP/dphi, P*m/cos in the synthesis and analysis layouts
        return {"basis": sixth, "analysis": sixth, "derivative": 2 * sixth,
                "zonal": 2 * sixth, "vector_fp64": low, "scratch": 0}

    # ---- validation -----------------------------------------------------
    def _grid_tensor(self, field):
        f = self.backend.asarray(field, dtype=self.backend.float_dtype).contiguous()
        if tuple(f.shape[-2:]) != self.grid.shape:
            raise ValueError(
                f"grid field has trailing shape {tuple(f.shape[-2:])}, expected {self.grid.shape}")
        return f

    def _spectral_tensor(self, coeff):
        c = self.backend.asarray(coeff, dtype=self.backend.complex_dtype).contiguous()
        if tuple(c.shape[-2:]) != self.spectral_shape:
            raise ValueError(
                f"spectral field has trailing shape {tuple(c.shape[-2:])}, "
                f"expected {self.spectral_shape}")
        return c

    @staticmethod
    def _count(t) -> int:
        return int(np.prod(t.shape[:-2], dtype=np.int64)) if t.dim() > 2 else 1

    def _empty_grid(self, lead):
        torch = self.backend.xp
        return torch.empty((*lead, *self.grid.shape), dtype=self.backend.float_dtype,
                           device=self.backend.device())

    def _empty_spectral(self, lead):
        torch = self.backend.xp
        return torch.empty((*lead, *self.spectral_shape), dtype=self.backend.complex_dtype,
                           device=self.backend.device())

    # ---- transforms -----------------------------------------------------
    def forward(self, field):
        """Analysis, grid -> spectral (transform.py:1486-1490)."""
        f = self._grid_tensor(field)
        out = self._empty_spectral(f.shape[:-2])
        _native.check(self._lib.arwen_sht_analysis(self._plan, _vp(f), _vp(out), self._count(f),
                                                   self.backend.stream()))
        return out

    def _synth(self, fn, coeff):
        c = self._spectral_tensor(coeff)
        out = self._empty_grid(c.shape[:-2])
        _native.check(fn(self._plan, _vp(c), _vp(out), self._count(c), self.backend.stream()))
        return out

    def inverse(self, coeff):
        """Synthesis, spectral -> grid (transform.py:1742-1746)."""
        return self._synth(self._lib.arwen_sht_synthesis, coeff)

    def inverse_zonal_derivative(self, coeff):
        """inverse(c * i m) (transform.py:1752-1760)."""
        return self._synth(self._lib.arwen_sht_zonal_derivative, coeff)

    def inverse_meridional_derivative(self, coeff):
        """d field / d latitude in grid space (transform.py:1748-1750)."""
        return self._synth(self._lib.arwen_sht_meridional_derivative, coeff)

    def gradient(self, coeff):
        """(east, north) derivatives per metre (transform.py:1762-1768)."""
        c = self._spectral_tensor(coeff)
        east = self._empty_grid(c.shape[:-2])
        north = self._empty_grid(c.shape[:-2])
        _native.check(self._lib.arwen_sht_gradient(self._plan, _vp(c), _vp(east), _vp(north),
                                                   self._count(c), self.backend.stream()))
        return east, north

    def _offset_tensor(self, value, count, dtype, what):
        """A per-field offset as a contiguous device array of ``count``
        entries: a scalar (every field), a sequence or an array."""
        if value is None:
            return None
        torch = self.backend.xp
        t = torch.as_tensor(value, dtype=dtype, device=self.backend.device())
        if t.dim() == 0:
            t = t.expand(count)
        t = t.reshape(-1).contiguous()
        if t.numel() != count:
            raise ValueError(f"{what} has {t.numel()} entries for {count} fields")
        return t

    def forward_about(self, field, grid_offset=None, coefficient_offset=None):
        """``forward(field - grid_offset) + coefficient_offset`` at (0, 0), one
        offset per field (leading dims flattened) or one for all: the theta
        analysis about whole-kelvin offsets of the reference
        (semilag/step.py:128-163, 610-621) and ``lnps_from_anomaly_coefficients``
        after an analysis (lnps_reference.py:47-49).  The subtraction happens
        in the float32 fold ahead of the FFT and the addition in the unpack,
        the same float32 operations the reference makes on the arrays."""
        f = self._grid_tensor(field)
        n = self._count(f)
        g = self._offset_tensor(grid_offset, n, self.backend.float_dtype, "grid_offset")
        c = self._offset_tensor(coefficient_offset, n, self.backend.complex_dtype,
                                "coefficient_offset")
        out = self._empty_spectral(f.shape[:-2])
        _native.check(self._lib.arwen_sht_analysis_about(
            self._plan, _vp(f), _vp(out), n, _vp(g), _vp(c), self.backend.stream()))
        return out

    def inverse_about(self, coeff, coefficient_offset=None, *, exp: bool = False):
        """``inverse(coeff + coefficient_offset at (0, 0))``, and its exp with
        ``exp=True``: ``lnps_from_anomaly_coefficients`` ahead of the inverse
        and the surface pressure ``exp(ln ps)`` the dycore takes from it
        (reference lnps_reference.py:47-49, dynamics.py:964-980)."""
        c = self._spectral_tensor(coeff)
        n = self._count(c)
        off = self._offset_tensor(coefficient_offset, n, self.backend.complex_dtype,
                                  "coefficient_offset")
        out = self._empty_grid(c.shape[:-2])
        _native.check(self._lib.arwen_sht_synthesis_about(
            self._plan, _vp(c), _vp(out), n, _vp(off), _native.JOB_EXP if exp else 0,
            self.backend.stream()))
        return out

    def inverse_with_gradient(self, coeff):
        """``(inverse(coeff), *gradient(coeff))`` from one pack and two plane
        contractions (the P and dP/dphi tables); the dycore takes both of
        temperature and ln ps every step (reference dynamics.py:907-1047)."""
        c = self._spectral_tensor(coeff)
        outs = [self._empty_grid(c.shape[:-2]) for _ in range(3)]
        _native.check(self._lib.arwen_sht_synthesis_gradient(
            self._plan, _vp(c), *(_vp(o) for o in outs), self._count(c),
            self.backend.stream()))
        return tuple(outs)

    def _spectral_op(self, op, coeff):
        c = self._spectral_tensor(coeff)
        out = self._empty_spectral(c.shape[:-2])
        _native.check(self._lib.arwen_sht_spectral_op(self._plan, op, _vp(c), _vp(out),
                                                      self._count(c), self.backend.stream()))
        return out

    def project(self, coeff):
        """Zero above the triangle, m = 0 real; a fresh array (transform.py:1124-1144)."""
        return self._spectral_op(_native.SPECTRAL_PROJECT, coeff)

    def laplacian(self, coeff):
        return self._spectral_op(_native.SPECTRAL_LAPLACIAN, coeff)

    def inverse_laplacian(self, coeff):
        return self._spectral_op(_native.SPECTRAL_INVERSE_LAPLACIAN, coeff)

    # ---- the Fourier waist (transform.py:1276-1472, 1629-1796) ----------
    def _edges(self, bands):
        return latitude_band_edges(self.grid.nlat,
                                   self.latitude_bands if bands is None else bands)

    def _waist_tensor(self, values):
        v = self.backend.asarray(values, dtype=self.backend.complex_dtype).contiguous()
        want = (self.grid.nlat, self.truncation + 1)
        if tuple(v.shape[-2:]) != want:
            raise ValueError(f"Fourier waist has trailing shape {tuple(v.shape[-2:])}, "
                             f"expected {want}")
        return v

    def _empty_waist(self, lead):
        torch = self.backend.xp
        return torch.empty((*lead, self.grid.nlat, self.truncation + 1),
                           dtype=self.backend.complex_dtype, device=self.backend.device())

    def _waist_from_rows(self, rows, values, row0: int, nrows: int) -> None:
        nlead = int(np.prod(rows.shape[:-2], dtype=np.int64))
        _native.check(self._lib.arwen_sht_waist_from_rows(
            self._plan, _vp(rows), _vp(values), nlead, int(nrows), int(row0),
            self.backend.stream()))

    def open_waist(self, lead, *, bands: int | None = None) -> FourierWaist:
        """An empty ``(*lead, nlat, T+1)`` waist for a band pipeline to fill
        (transform.py:1276-1292)."""
        return FourierWaist.open(self, lead, self._edges(bands))

    def fourier_waist(self, field, *, bands: int | None = None) -> FourierWaist:
        """The head of an analysis: rfft of every latitude row / nlon, orders
        m <= T (transform.py:1294-1404).  Each row is transformed alone, so
        the band count changes no bit; the waist carries the schedule."""
        f = self._grid_tensor(field)
        edges = self._edges(bands)
        values = self._empty_waist(f.shape[:-2])
        self._waist_from_rows(f, values, 0, self.grid.nlat)
        return FourierWaist(self, values, edges)

    def _require(self, table, allowed, what):
        if table is None:
            return allowed[0]
        if not isinstance(table, LegendreTable) or table not in allowed:
            names = ", ".join(t.name for t in allowed)
            raise ValueError(f"{what} contracts with the plan's {names} table; got {table!r}")
        return table

    def _analysis_from_waist(self, values):
        out = self._empty_spectral(values.shape[:-2])
        _native.check(self._lib.arwen_sht_analysis_from_waist(
            self._plan, _vp(values), _vp(out), self._count(values), self.backend.stream()))
        return out

    def contract_waist(self, waist: FourierWaist, table=None):
        """The tail of an analysis: contract a full waist to coefficients,
        m = 0 realified; consumes the waist (transform.py:1435-1452)."""
        self._require(table, (self._analysis,), "contract_waist")
        waist.close()
        return self._analysis_from_waist(self._waist_tensor(waist.take()))

    def contract_values(self, values, table=None):
        """:meth:`contract_waist` on a waist's buffer or a leading slice of it
        (transform.py:1454-1472)."""
        self._require(table, (self._analysis,), "contract_values")
        return self._analysis_from_waist(self._waist_tensor(values))

    def _to_waist(self, coeff, kind: int):
        c = self._spectral_tensor(coeff)
        values = self._empty_waist(c.shape[:-2])
        _native.check(self._lib.arwen_sht_synthesis_to_waist(
            self._plan, _vp(c), _vp(values), self._count(c), int(kind), self.backend.stream()))
        return values

    def contract_to_waist(self, coeff, table=None, *, bands: int | None = None) -> FourierWaist:
        """The head of a synthesis: the Legendre sum of ``coeff`` with the
        basis (or dP/dphi) into a waist that carries the drain schedule
        (transform.py:1629-1652)."""
        table = self._require(table, (self._basis, self._derivative_basis), "contract_to_waist")
        kind = _native.WAIST_BASIS if table is self._basis else _native.WAIST_DERIVATIVE
        return FourierWaist(self, self._to_waist(coeff, kind), self._edges(bands))

    def _rows_to_grid(self, values, waist: FourierWaist, r0: int, r1: int):
        r0, r1 = int(r0), int(r1)
        if not (0 <= r0 < r1 <= self.grid.nlat):
            raise ValueError(f"latitude band [{r0}, {r1}) is not a nonempty range of the "
                             f"{self.grid.nlat} rows")
        torch = self.backend.xp
        out = torch.empty((*waist.lead, r1 - r0, self.grid.nlon),
                          dtype=self.backend.float_dtype, device=self.backend.device())
        nlead = int(np.prod(waist.lead, dtype=np.int64))
        _native.check(self._lib.arwen_sht_rows_from_waist(
            self._plan, _vp(values), _vp(out), nlead, r1 - r0, r0, self.backend.stream()))
        return out

    def waist_band_to_grid(self, waist: FourierWaist, r0: int, r1: int):
        """One latitude band of a waist in grid space, the waist kept
        (transform.py:1704-1711)."""
        return self._rows_to_grid(waist.values, waist, r0, r1)

    def waist_to_grid(self, waist: FourierWaist):
        """The whole waist in grid space; consumes it (transform.py:1713-1735).
        One call drains every band: each row's inverse FFT is its own, so
        the result is the band-by-band drain's."""
        values = waist.values
        out = self._rows_to_grid(values, waist, 0, self.grid.nlat)
        waist.take()
        return out

    def gradient_waists(self, coeff, *, bands: int | None = None):
        """The two contractions of :meth:`gradient` before their drains:
        (c i m against the basis, c against dP/dphi) (transform.py:1770-1787)."""
        c = self._spectral_tensor(coeff)
        edges = self._edges(bands)
        return (FourierWaist(self, self._to_waist(c, _native.WAIST_ZONAL), edges),
                FourierWaist(self, self._to_waist(c, _native.WAIST_DERIVATIVE), edges))

    def gradient_band(self, waists, r0: int, r1: int):
        """One latitude band of :meth:`gradient` from its waists
        (transform.py:1789-1796)."""
        east_waist, north_waist = waists
        east = self.waist_band_to_grid(east_waist, r0, r1)
        east /= self.grid.radius_m * self._coslat[int(r0):int(r1)]
        north = self.waist_band_to_grid(north_waist, r0, r1)
        north /= self.grid.radius_m
        return east, north

    # ---- small helpers (transform.py:1813-1835) -------------------------
    def zeros(self, *leading: int):
        torch = self.backend.xp
        return torch.zeros((*leading, *self.spectral_shape), dtype=self.backend.complex_dtype,
                           device=self.backend.device())

    def constant_coeff(self, value: float):
        out = self.zeros()
        out[0, 0] = float(value) * math.sqrt(4.0 * math.pi)
        return out

    def add_grid_constant(self, coeff, value: float):
        out = self._spectral_tensor(coeff).clone()
        out[..., 0, 0] += float(value) * math.sqrt(4.0 * math.pi)
        return out

    def spectral_mean_square(self, coeff) -> float:
        c = self.backend.to_numpy(coeff)
        if tuple(c.shape[-2:]) != self.spectral_shape:
            raise ValueError(f"spectral field has trailing shape {c.shape[-2:]}, "
                             f"expected {self.spectral_shape}")
        w = np.ones(self.truncation + 1)
        w[1:] = 2.0
        total = np.sum(np.abs(c) ** 2 * w[None, :], axis=(-2, -1))
        return float(np.asarray(total).mean() / (4.0 * np.pi))

    def power_by_degree(self, coeff) -> np.ndarray:
        c = self.backend.to_numpy(coeff)
        w = np.ones(self.truncation + 1)
        w[1:] = 2.0
        return np.sum(np.abs(c) ** 2 * w[None, :], axis=-1)

    def transform_check(self, *, seed: int = 0, max_degree: int | None = None) -> dict:
        """Round-trip and Parseval gate (reference transform.py:1837-1863)."""
        rng = np.random.default_rng(seed)
        t = self.truncation if max_degree is None else min(self.truncation, int(max_degree))
        coeff = np.zeros(self.spectral_shape, dtype=np.complex128)
        for n in range(t + 1):
            for m in range(n + 1):
                coeff[n, m] = rng.normal() + (0.0j if m == 0 else 1j * rng.normal())
        grid = self.inverse(coeff)
        back = self.backend.to_numpy(self.forward(grid))
        scale = max(1.0, float(np.max(np.abs(coeff))))
        error = float(np.max(np.abs(back - coeff)) / scale)
        spatial = self.grid.global_mean(self.backend.to_numpy(grid).astype(np.float64) ** 2)
        spectral = float(np.sum(np.abs(coeff) ** 2 * np.where(
            np.arange(self.truncation + 1)[None, :] == 0, 1.0, 2.0)) / (4.0 * np.pi))
        parseval = abs(spatial - spectral) / max(1.0, abs(spectral))
        return {"truncation": self.truncation, "nlat": self.grid.nlat, "nlon": self.grid.nlon,
                "roundtrip_relative_linf": error, "parseval_relative_error": float(parseval)}
