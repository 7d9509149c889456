"""Spectral scalars, gradients and winds at arbitrary latitude/longitude points.

Drop-in for the reference's ``arwen_global.spectral.sampling``
(src/arwen_global/spectral/sampling.py): same functions, arguments, error
conditions and result layout ``(*lead, *point_shape)`` float64.  Every call
is one pack and one launch of ``libarwen_sht.so``'s point sampler: one lane
per point runs the FP64 normalized recurrence (the reference's operand
order) and multiply-adds a tile of fields' coefficients into per-order
amplitudes; nothing the size of the reference's (2K x points) basis is
materialised, so ``chunk_points`` only bounds a launch and never the
arithmetic of a point.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .transform import SphericalHarmonicTransform

#: The reference's host-path workspace bound (sampling.py:22-26); kept for
#: callers that read it.  This engine has no basis workspace to bound.
DEFAULT_SAMPLE_WORKSPACE_BYTES = 64 * 1024**2
#: The reference's device-path workspace bound (sampling.py:28-33).  Here the
#: workspace holds the packed coefficients and per-part sums per point; with
#: ``chunk_points=None`` the point chunk is the largest that keeps it within
#: this bound (the cuts depend on T only, so chunking never changes a bit).
DEFAULT_DEVICE_SAMPLE_WORKSPACE_BYTES = 512 * 1024**2
#: ``"float64"``: basis and sums in FP64.  ``"float32"``: each basis value
#: rounded to fp32 and one order's terms summed in fp32, the orders summed in
#: FP64 (the reference's state-precision sampler, sampling.py:35-44).
SAMPLE_PRECISIONS = ("float64", "float32")
#: The reference's float32 GEMM block (sampling.py:47); here the fp32 partial
#: sums are per order, at most T+1 terms.
STATE_PRECISION_BLOCK = 256

_KIND = {"scalar": 0, "gradient": 1, "wind": 2}


def _resolve_precision(precision) -> str:
    if precision is None:
        return "float64"
    if precision not in SAMPLE_PRECISIONS:
        raise ValueError(f"sample precision must be one of {SAMPLE_PRECISIONS}, got {precision!r}")
    return str(precision)


def triangle_index(truncation: int) -> tuple[np.ndarray, np.ndarray]:
    """``(n, m)`` of every coefficient inside the triangle in the packed,
    order-major order (``m = 0..T``, ``n = m..T`` inside each order) the
    sampler walks (reference sampling.py:72-80)."""
    t = int(truncation)
    ms = np.concatenate([np.full(t + 1 - m, m, dtype=np.int64) for m in range(t + 1)])
    ns = np.concatenate([np.arange(m, t + 1, dtype=np.int64) for m in range(t + 1)])
    return ns, ms


def _points(latitude_deg, longitude_deg):
    """Flattened radians and the point shape (reference sampling.py:346-356)."""
    lat, lon = np.broadcast_arrays(np.asarray(latitude_deg, dtype=np.float64),
                                   np.asarray(longitude_deg, dtype=np.float64))
    if not np.isfinite(lat).all() or not np.isfinite(lon).all():
        raise ValueError("sample coordinates must be finite")
    if np.any((lat < -90.0) | (lat > 90.0)):
        raise ValueError("latitude must lie in [-90, 90] degrees")
    return np.deg2rad(lat.ravel()), np.deg2rad(lon.ravel()), lat.shape


def _chunk(requested, npts: int, fixed_bytes: int = 0, per_point_bytes: int = 0) -> int:
    if requested is None:
        if per_point_bytes <= 0:
            return max(1, npts)
        room = max(DEFAULT_DEVICE_SAMPLE_WORKSPACE_BYTES - fixed_bytes, per_point_bytes)
        return max(1, min(npts, room // per_point_bytes))
    if isinstance(requested, bool) or not isinstance(requested, (int, np.integer)) \
            or int(requested) < 1:
        raise ValueError("chunk_points must be a positive integer")
    return int(requested)


def _sample(transform: SphericalHarmonicTransform, kind: str, first, second, latitude_deg,
            longitude_deg, chunk_points, precision, as_device):
    prec = _resolve_precision(precision)
    t = transform
    torch = t.backend.xp
    dev = t.backend.device()
    c0 = t._spectral_tensor(first)
    c1 = None
    if kind == "wind":
        c1 = t._spectral_tensor(second)
        if c1.shape != c0.shape:
            raise ValueError(f"vorticity/divergence shape mismatch: {tuple(c0.shape)} != "
                             f"{tuple(c1.shape)}")
    lat, lon, point_shape = _points(latitude_deg, longitude_deg)
    coslat = np.cos(lat)
    if kind != "scalar" and np.any(np.abs(coslat) < 1.0e-12):
        raise ValueError("vector/gradient sampling is undefined at the exact poles")
    lead = tuple(c0.shape[:-2])
    count = t._count(c0)
    lib = t._lib
    code = _KIND[kind]
    fixed = int(lib.arwen_sht_sample_workspace_bytes(t._plan, count, code, 0))
    per_point = int(lib.arwen_sht_sample_workspace_bytes(t._plan, count, code, 1)) - fixed
    chunk = _chunk(chunk_points, lat.size, fixed, per_point)
    npts = int(lat.size)
    f64 = torch.float64
    outs = [torch.empty((count, npts), dtype=f64, device=dev)
            for _ in range(1 if kind == "scalar" else 2)]
    if count and npts:
        x = torch.from_numpy(np.sin(lat)).to(dev)
        cl = torch.from_numpy(coslat).to(dev)
        lo = torch.from_numpy(lon).to(dev)
        nbytes = int(lib.arwen_sht_sample_workspace_bytes(t._plan, count, code,
                                                          min(chunk, npts)))
        work = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        vp = ctypes.c_void_p
        for start in range(0, npts, chunk):
            n = min(chunk, npts - start)
            # a chunk's outputs are columns start.. of row-major (count, npts):
            # evaluate into contiguous (count, n) and copy them in
            part = [torch.empty((count, n), dtype=f64, device=dev) for _ in outs] \
                if n != npts else outs
            _native.check(lib.arwen_sht_sample(
                t._plan, vp(c0.data_ptr()), vp(c1.data_ptr() if c1 is not None else 0), count,
                vp(x[start:].data_ptr()), vp(cl[start:].data_ptr()), vp(lo[start:].data_ptr()),
                n, code, 0 if prec == "float64" else 1, vp(work.data_ptr()), nbytes,
                vp(part[0].data_ptr()), vp(part[1].data_ptr() if len(part) > 1 else 0),
                t.backend.stream()))
            if part is not outs:
                for o, p in zip(outs, part):
                    o[:, start:start + n] = p
    shaped = [o.reshape((*lead, *point_shape)) for o in outs]
    if not as_device:
        shaped = [t.backend.to_numpy(o) for o in shaped]
    return shaped[0] if kind == "scalar" else (shaped[0], shaped[1])


def sample_scalar(transform: SphericalHarmonicTransform, coeff, latitude_deg, longitude_deg, *,
                  chunk_points: int | None = None, precision: str | None = None,
                  as_device: bool = False):
    """Real scalar fields at arbitrary points, leading dimensions kept; exact
    poles admitted (reference sampling.py:382-426).  ``as_device`` returns
    the device tensor instead of numpy."""
    return _sample(transform, "scalar", coeff, None, latitude_deg, longitude_deg,
                   chunk_points, precision, as_device)


def sample_gradient(transform: SphericalHarmonicTransform, coeff, latitude_deg, longitude_deg, *,
                    chunk_points: int | None = None, precision: str | None = None,
                    as_device: bool = False):
    """Eastward and northward derivatives per metre at non-polar points
    (reference sampling.py:429-478)."""
    return _sample(transform, "gradient", coeff, None, latitude_deg, longitude_deg,
                   chunk_points, precision, as_device)


def sample_wind(transform: SphericalHarmonicTransform, vorticity, divergence, latitude_deg,
                longitude_deg, *, chunk_points: int | None = None, precision: str | None = None,
                as_device: bool = False):
    """Eastward and northward wind from vorticity and divergence, through
    psi, chi = inverse_laplacian(...) and their gradients, in one launch
    (reference sampling.py:481-515)."""
    return _sample(transform, "wind", vorticity, divergence, latitude_deg, longitude_deg,
                   chunk_points, precision, as_device)


def regular_latlon_coordinates(nlat: int, nlon: int, *, include_poles: bool = False):
    """Cell-centre regular global coordinates, or pole-inclusive latitudes
    (reference sampling.py:518-537)."""
    if isinstance(nlat, bool) or isinstance(nlon, bool):
        raise ValueError("regular-grid dimensions must be integers")
    ny, nx = int(nlat), int(nlon)
    if ny != nlat or nx != nlon:
        raise ValueError("regular-grid dimensions must be integers")
    if ny < 2 or nx < 4:
        raise ValueError("regular grid needs nlat>=2 and nlon>=4")
    if include_poles:
        lat = np.linspace(-90.0, 90.0, ny, dtype=np.float64)
    else:
        lat = -90.0 + (np.arange(ny, dtype=np.float64) + 0.5) * 180.0 / ny
    lon = np.arange(nx, dtype=np.float64) * 360.0 / nx
    return lat, lon


def regular_latlon_scalar(transform: SphericalHarmonicTransform, coeff, *, nlat: int, nlon: int,
                          include_poles: bool = False, chunk_points: int | None = None):
    """``(lat, lon, field)`` of a scalar on a regular grid (reference
    sampling.py:540-556)."""
    lat, lon = regular_latlon_coordinates(nlat, nlon, include_poles=include_poles)
    lon2, lat2 = np.meshgrid(lon, lat)
    return lat, lon, sample_scalar(transform, coeff, lat2, lon2, chunk_points=chunk_points)


__all__ = [
    "DEFAULT_DEVICE_SAMPLE_WORKSPACE_BYTES", "DEFAULT_SAMPLE_WORKSPACE_BYTES",
    "SAMPLE_PRECISIONS", "STATE_PRECISION_BLOCK", "regular_latlon_coordinates",
    "regular_latlon_scalar", "sample_gradient", "sample_scalar", "sample_wind",
    "triangle_index",
]
