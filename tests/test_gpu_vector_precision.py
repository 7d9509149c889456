"""The vector operators' FP64 low orders against the FP64 oracle (needs an MI355X).

Vorticity and divergence at high degree are small differences of the wind's
large low-order values, weighted by dP/dphi (which grows like n), so fp32
rounding on the low orders is amplified by ~T/m.  With red-spectrum inputs
at the benchmark's class scales (vorticity 1e-5, divergence 2e-6 s^-1) the
plain split-BF16 and FP32 paths leave 4e-6 / 2e-5 (T383) in the round trip.
The split-BF16 plans therefore run the orders m < 4 ceil((T+1)/192) of
uv_from_vordiv and vordiv_from_uv in FP64 end to end (FP64 FFTs of u and v,
FP64 tables and sums); these tests hold that path to the 1e-6 bound of the
north star, measured as max|c - c_fp64| / max|c_fp64| per field class.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle.sht import OracleTransform, red_spectrum

from _helpers import rel_err

pytestmark = pytest.mark.gpu

BOUND = 1e-6  # relative error per field class vs the FP64 oracle


def _transform(T, precision="bf16x6", fp64_orders=None):
    from arwen_spectral import SphericalHarmonicTransform

    return SphericalHarmonicTransform.create(T, dealias_factor=1.0, backend="hip",
                                             precision=precision, fp64_orders=fp64_orders)


def _vector_fields(T, levels, seed):
    zeta = red_spectrum(T, (levels,), seed=seed, scale=1e-5)
    div = red_spectrum(T, (levels,), seed=seed + 1, scale=2e-6)
    zeta[..., 0, 0] = 0.0
    div[..., 0, 0] = 0.0
    return zeta, div


@pytest.fixture(scope="module")
def oracle383():
    return OracleTransform.create(383, dealias_factor=1.0)


@pytest.mark.parametrize("T", [63, 127, 383])
def test_vector_round_trip_within_the_bound(T, oracle383):
    from arwen_spectral import VorticityDivergenceOperator

    o = oracle383 if T == 383 else OracleTransform.create(T, dealias_factor=1.0)
    tr = _transform(T)
    assert tr.vector_fp64_orders == 4 * ((T + 1 + 191) // 192)
    op = VorticityDivergenceOperator(tr)
    zeta, div = _vector_fields(T, 4, seed=30 + T)
    u64, v64 = o.wind_from_vordiv(zeta, div)
    z64, d64 = o.vordiv_from_wind(u64, v64)
    u, v = op.wind_from_vordiv(zeta, div)
    z, d = op.vordiv_from_wind(u, v)
    errs = {"u": rel_err(u, u64), "v": rel_err(v, v64),
            "vorticity": rel_err(z, z64), "divergence": rel_err(d, d64)}
    print(f"T{T} vector round trip vs FP64: " + ", ".join(f"{k} {e:.2e}" for k, e in errs.items()))
    for k, e in errs.items():
        assert e < BOUND, (k, e)
    # analysis alone, of the oracle's winds rounded to the boundary's fp32
    u32, v32 = u64.astype(np.float32), v64.astype(np.float32)
    za, da = op.vordiv_from_wind(u32, v32)
    zr, dr = o.vordiv_from_wind(u32.astype(np.float64), v32.astype(np.float64))
    assert max(rel_err(za, zr), rel_err(da, dr)) < BOUND


@pytest.mark.parametrize("T,nlat,nlon,fp64", [(94, 96, 192, None), (62, 72, 128, 12),
                                               (100, 102, 256, 36)])
def test_recurrence_form_on_uneven_shapes(T, nlat, nlon, fp64):
    """The orders m >= m0 run in the recurrence form (the basis to degree
    T+1, pack and unpack with halo degrees): truncations whose T+1 and T+2
    fall on and off the kernels' 64-degree chunk edges, latitudes padded past
    the truncation, and m0 inside the first 32-order tile."""
    from arwen_spectral import SphericalHarmonicTransform, VorticityDivergenceOperator

    tr = SphericalHarmonicTransform.create(T, nlat=nlat, nlon=nlon, dealias_factor=1.0,
                                           backend="hip", precision="bf16x6",
                                           fp64_orders=fp64)
    assert tr.vector_fp64_orders == (4 if fp64 is None else fp64)
    o = OracleTransform.create(T, nlat=nlat, nlon=nlon, dealias_factor=1.0)
    op = VorticityDivergenceOperator(tr)
    zeta, div = _vector_fields(T, 3, seed=200 + T)
    u64, v64 = o.wind_from_vordiv(zeta, div)
    u, v = op.wind_from_vordiv(zeta, div)
    assert max(rel_err(u, u64), rel_err(v, v64)) < BOUND
    uh = u.cpu().numpy().astype(np.float64)
    vh = v.cpu().numpy().astype(np.float64)
    z64, d64 = o.vordiv_from_wind(uh, vh)
    z, d = op.vordiv_from_wind(u, v)
    print(f"T{T} {nlat}x{nlon} m0={tr.vector_fp64_orders}: analysis vorticity "
          f"{rel_err(z, z64):.2e}, divergence {rel_err(d, d64):.2e}")
    assert max(rel_err(z, z64), rel_err(d, d64)) < BOUND


@pytest.mark.parametrize("T,nlat,nlon", [(94, 96, 192), (62, 72, 128), (100, 102, 256)])
def test_plain_fp32_recurrence_form_on_uneven_shapes(T, nlat, nlon):
    """The plain FP32 path (no FP64 orders) on wave-resident grids: orders
    m >= 4 in the recurrence form, the first order tile against the P m/cos
    and dP/dphi tables, the fold and unfold scaling the two apart.  Same
    checks and bound as the T21/T63 vector parity tests."""
    from arwen_spectral import SphericalHarmonicTransform, VorticityDivergenceOperator

    tr = SphericalHarmonicTransform.create(T, nlat=nlat, nlon=nlon, dealias_factor=1.0,
                                           backend="hip", precision="float32")
    assert tr.vector_fp64_orders == 0
    o = OracleTransform.create(T, nlat=nlat, nlon=nlon, dealias_factor=1.0)
    op = VorticityDivergenceOperator(tr)
    zeta = o.project(red_spectrum(T, (2,), seed=300 + T, scale=1e-5))
    div = o.project(red_spectrum(T, (2,), seed=400 + T, scale=1e-5))
    u, v = op.wind_from_vordiv(zeta, div)
    u64, v64 = o.wind_from_vordiv(zeta, div)
    assert max(rel_err(u, u64), rel_err(v, v64)) < BOUND
    rng = np.random.default_rng(T)
    gu = rng.standard_normal((2, nlat, nlon)) * 10
    gv = rng.standard_normal((2, nlat, nlon)) * 10
    z, d = op.vordiv_from_wind(gu, gv)
    z64, d64 = o.vordiv_from_wind(gu, gv)
    print(f"FP32 T{T} {nlat}x{nlon}: winds {max(rel_err(u, u64), rel_err(v, v64)):.2e}, "
          f"vorticity {rel_err(z, z64):.2e}, divergence {rel_err(d, d64):.2e}")
    assert max(rel_err(z, z64), rel_err(d, d64)) < BOUND


def test_the_fp64_orders_carry_the_bound(oracle383):
    """Without the FP64 orders the same round trip misses the bound by an
    order of magnitude (the reason they exist); with them, split-BF16 and
    FP32 at m >= m0 land in the same class."""
    from arwen_spectral import VorticityDivergenceOperator

    o = oracle383
    zeta, div = _vector_fields(383, 2, seed=77)
    z64, d64 = o.vordiv_from_wind(*o.wind_from_vordiv(zeta, div))
    errs = {}
    for precision, m0 in (("bf16x6", 0), ("bf16x6", None), ("float32", 8), ("float32", None)):
        tr = _transform(383, precision, m0)
        op = VorticityDivergenceOperator(tr)
        z, d = op.vordiv_from_wind(*op.wind_from_vordiv(zeta, div))
        errs[(precision, tr.vector_fp64_orders)] = max(rel_err(z, z64), rel_err(d, d64))
    print("T383 vector round trip (precision, fp64 orders) -> error:", errs)
    assert errs[("bf16x6", 0)] > 10 * BOUND
    assert errs[("float32", 0)] > 10 * BOUND
    assert errs[("bf16x6", 8)] < BOUND
    assert errs[("float32", 8)] < BOUND


def test_fp64_orders_resolution_and_identity():
    from arwen_spectral import SphericalHarmonicTransform

    assert _transform(63, "float32").vector_fp64_orders == 0
    assert _transform(63, "bf16x6").vector_fp64_orders == 4
    assert _transform(63, "bf16x6", 5).vector_fp64_orders == 8
    assert _transform(63, "bf16x6", 0).vector_fp64_orders == 0
    assert _transform(63, "float32", 200).vector_fp64_orders == 64  # all orders
    # 33 x 66 (T21's default grid): nlon is not 64 L, so no FP64 FFT and no FP64 orders
    assert SphericalHarmonicTransform.create(21, precision="bf16x6").vector_fp64_orders == 0
    with pytest.raises(ValueError, match="fp64_orders"):
        _transform(63, "bf16x6", -2)
    plain, split = _transform(63, "float32"), _transform(63, "bf16x6")
    assert "vector_fp64_orders" not in plain.identity
    assert split.identity["vector_fp64_orders"] == 4
    assert _transform(63, "bf16x6", 0).identity_hash != split.identity_hash


@pytest.mark.parametrize("T", [63, 383])
def test_all_orders_in_fp64(T, oracle383):
    """fp64_orders above T runs every order through the FP64 kernels (the
    grouped GEMM sees no vector work): the vector operators then match the
    oracle to the fp32 rounding of their inputs and outputs.  The analysis is
    compared with the oracle's analysis of the same fp32 winds: rounding the
    exact winds to fp32 alone moves this seed's divergence by 9.3e-7 at T383
    (the boundary's floor, which the other tests' seeds keep near 2e-7)."""
    from arwen_spectral import VorticityDivergenceOperator

    o = oracle383 if T == 383 else OracleTransform.create(T, dealias_factor=1.0)
    tr = _transform(T, "bf16x6", T + 1)
    assert tr.vector_fp64_orders == 4 * ((T + 4) // 4)
    op = VorticityDivergenceOperator(tr)
    zeta, div = _vector_fields(T, 2, seed=50 + T)
    u64, v64 = o.wind_from_vordiv(zeta, div)
    u, v = op.wind_from_vordiv(zeta, div)
    assert max(rel_err(u, u64), rel_err(v, v64)) < 2e-7
    uh = u.cpu().numpy().astype(np.float64)
    vh = v.cpu().numpy().astype(np.float64)
    z64, d64 = o.vordiv_from_wind(uh, vh)
    z, d = op.vordiv_from_wind(u, v)
    assert max(rel_err(z, z64), rel_err(d, d64)) < 1e-7


def test_t767_vector_round_trip_within_the_bound():
    """T767 on 768 x 1536 (nlon = 64 x 24, the longest wave-resident FFT;
    16 FP64 orders): the vector round trip against its inputs."""
    from arwen_spectral import VorticityDivergenceOperator

    tr = _transform(767)
    assert tr.vector_fp64_orders == 16
    op = VorticityDivergenceOperator(tr)
    zeta, div = _vector_fields(767, 2, seed=90)
    z, d = op.vordiv_from_wind(*op.wind_from_vordiv(zeta, div))
    ez, ed = rel_err(z, zeta), rel_err(d, div)
    print(f"T767 vector round trip: vorticity {ez:.2e}, divergence {ed:.2e}")
    assert max(ez, ed) < BOUND


def test_identity_rules_follow_the_reference():
    """The checkpoint contract (reference transform.py:1056-1104 and its
    tests test_arwen_global_memory_levers.py:135-149,
    test_global_spectral_fourier_waist.py:355-374): the float32 payload has
    the reference's keys (geometry, backend, float_dtype); arithmetic that
    moves bits joins (a non-default legendre_band, a split precision, FP64
    vector orders); latitude_bands, which moves no bits, never joins; the
    hashes are SHA-256 of the payload's sorted compact JSON."""
    import hashlib
    import json

    from arwen_spectral import SphericalHarmonicTransform

    base = _transform(63, "float32")
    assert set(base.identity) == {"truncation", "nlat", "nlon", "radius_m", "normalization",
                                  "latitude_grid", "backend", "float_dtype"}
    assert base.identity["float_dtype"] == "float32"
    assert base.geometry_identity["normalization"] == "complex-orthonormal-condon-shortley"
    assert base.geometry_identity["latitude_grid"] == "gauss-legendre-in-sin-latitude"
    canon = lambda d: hashlib.sha256(json.dumps(d, sort_keys=True, separators=(",", ":"))
                                     .encode()).hexdigest()
    assert base.identity_hash == canon(base.identity)
    assert base.geometry_hash == canon(base.geometry_identity)
    banded = SphericalHarmonicTransform.create(63, dealias_factor=1.0, precision="float32",
                                               latitude_bands=4)
    assert banded.identity == base.identity and banded.identity_hash == base.identity_hash
    moved = SphericalHarmonicTransform.create(63, dealias_factor=1.0, precision="float32",
                                              legendre_band=8)
    assert moved.identity["legendre_band"] == 8 and moved.identity_hash != base.identity_hash
    split = _transform(63, "bf16x6")
    assert split.identity["legendre_precision"] == "bf16x6"
    assert split.geometry_hash == base.geometry_hash
    assert split.identity_hash != base.identity_hash
    assert set(split.identity) - set(base.identity) == {"legendre_precision", "vector_fp64_orders"}
