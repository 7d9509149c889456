"""The point sampler on the host: the oracle's restatement of the reference's
host sampler (spectral/sampling.py:382-515) pinned against closed forms and
independent routes, and the engine's argument rules (no GPU needed).
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import sampling as osamp
from oracle.sht import OracleTransform, red_spectrum


def _random_points(rng, n, *, poles=False):
    lat = rng.uniform(-89.0, 89.0, n)
    lon = rng.uniform(-180.0, 540.0, n)
    if poles:
        lat[:2] = [90.0, -90.0]
    return lat, lon


def test_oracle_sampler_single_modes_match_scipy_sph_harm_y():
    """c_nm = a + ib alone: the field is mult * Re(c Y_nm) with scipy's
    orthonormal Condon-Shortley harmonic (polar angle = 90 deg - lat)."""
    from scipy.special import sph_harm_y

    o = OracleTransform.create(12)
    rng = np.random.default_rng(5)
    lat, lon = _random_points(rng, 40, poles=True)
    theta, phi = np.deg2rad(90.0 - lat), np.deg2rad(lon)
    for n, m in ((0, 0), (1, 0), (1, 1), (5, 3), (12, 12), (12, 1), (9, 0)):
        c = np.zeros((13, 13), dtype=np.complex128)
        c[n, m] = 0.7 - 0.4j if m else 0.7
        want = (1.0 if m == 0 else 2.0) * (c[n, m] * sph_harm_y(n, m, theta, phi)).real
        got = osamp.sample_scalar(o, c, lat, lon)
        assert np.max(np.abs(got - want)) < 1e-13, (n, m)


def test_oracle_sampler_at_grid_points_is_the_fft_synthesis():
    o = OracleTransform.create(21)
    c = red_spectrum(21, (2,), seed=6)
    lat = np.rad2deg(np.arcsin(o.grid.sin_lat))
    lon = np.rad2deg(o.grid.lon_rad)
    lon2, lat2 = np.meshgrid(lon, lat)
    f = osamp.sample_scalar(o, c, lat2, lon2)
    assert f.shape == (2, o.grid.nlat, o.grid.nlon)
    assert np.max(np.abs(f - o.inverse(c))) < 1e-12 * np.max(np.abs(f))
    east, north = osamp.sample_gradient(o, c, lat2, lon2)
    ge, gn = o.gradient(c)
    assert np.max(np.abs(east - ge)) < 1e-12 * np.max(np.abs(ge))
    assert np.max(np.abs(north - gn)) < 1e-12 * np.max(np.abs(gn))
    z, d = o.project(red_spectrum(21, (2,), seed=7)), o.project(red_spectrum(21, (2,), seed=8))
    u, v = osamp.sample_wind(o, z, d, lat2, lon2)
    wu, wv = o.wind_from_vordiv(z, d)
    assert np.max(np.abs(u - wu)) < 1e-12 * np.max(np.abs(wu))
    assert np.max(np.abs(v - wv)) < 1e-12 * np.max(np.abs(wv))


def test_oracle_sampler_gradient_matches_finite_differences_off_the_grid():
    o = OracleTransform.create(15)
    c = red_spectrum(15, (), seed=9)
    rng = np.random.default_rng(10)
    lat, lon = _random_points(rng, 25)
    h = 1e-4  # degrees
    east, north = osamp.sample_gradient(o, c, lat, lon)
    a = o.grid.radius_m
    d_lon = (osamp.sample_scalar(o, c, lat, lon + h) - osamp.sample_scalar(o, c, lat, lon - h)) \
        / (2 * np.deg2rad(h))
    d_lat = (osamp.sample_scalar(o, c, lat + h, lon) - osamp.sample_scalar(o, c, lat - h, lon)) \
        / (2 * np.deg2rad(h))
    scale = np.max(np.abs(d_lat)) / a
    assert np.max(np.abs(east - d_lon / (a * np.cos(np.deg2rad(lat))))) < 1e-6 * scale
    assert np.max(np.abs(north - d_lat / a)) < 1e-6 * scale


def test_sampler_argument_rules_without_a_device():
    from arwen_spectral import sampling

    assert sampling._resolve_precision(None) == "float64"
    with pytest.raises(ValueError, match="sample precision must be one of"):
        sampling._resolve_precision("float16")
    with pytest.raises(ValueError, match="must be finite"):
        sampling._points([np.nan], [0.0])
    with pytest.raises(ValueError, match=r"latitude must lie in \[-90, 90\]"):
        sampling._points([91.0], [0.0])
    lat, lon, shape = sampling._points([[10.0, 20.0]], 5.0)
    assert shape == (1, 2) and np.allclose(np.rad2deg(lon), 5.0)
    for bad in (0, -3, 2.5, True):
        with pytest.raises(ValueError, match="chunk_points must be a positive integer"):
            sampling._chunk(bad, 10)
    assert sampling._chunk(None, 0) == 1 and sampling._chunk(7, 10) == 7
    ns, ms = sampling.triangle_index(3)
    assert list(zip(ns, ms))[:5] == [(0, 0), (1, 0), (2, 0), (3, 0), (1, 1)]
    assert ns.size == 10
    la, lo = sampling.regular_latlon_coordinates(4, 8)
    assert np.allclose(la, [-67.5, -22.5, 22.5, 67.5]) and np.allclose(lo, np.arange(8) * 45.0)
    la, _ = sampling.regular_latlon_coordinates(3, 4, include_poles=True)
    assert np.allclose(la, [-90.0, 0.0, 90.0])
    for args in ((1, 8), (4, 3), (True, 8), (4.5, 8)):
        with pytest.raises(ValueError):
            sampling.regular_latlon_coordinates(*args)
