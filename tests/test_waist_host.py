"""The Fourier waist on the host: the latitude band schedule and the row
bookkeeping of a band fill (reference spectral/transform.py:68-94, 373-414),
and the oracle's waist maps pinned against closed forms.

Known answers for the band schedule are the reference's own
(tests/test_global_spectral_fourier_waist.py:67-95).
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle.sht import OracleTransform, red_spectrum


def test_the_band_schedule_tiles_the_latitudes_exactly_once():
    from arwen_spectral.waist import latitude_band_edges

    for nlat in (1, 2, 7, 32, 64, 801):
        for bands in (1, 2, 3, 5, 8, 32, 1000):
            edges = latitude_band_edges(nlat, bands)
            assert edges[0][0] == 0 and edges[-1][1] == nlat
            assert all(b > a for a, b in edges)
            assert all(edges[i][1] == edges[i + 1][0] for i in range(len(edges) - 1))
            assert len(edges) == min(bands, nlat)


def test_the_band_schedule_known_answer():
    from arwen_spectral.waist import latitude_band_edges

    assert latitude_band_edges(801, 8) == [(0, 100), (100, 200), (200, 300), (300, 400),
                                           (400, 500), (500, 600), (600, 700), (700, 801)]
    # floor(k n / B): uneven bands keep the remainder at the top of each cut
    assert latitude_band_edges(33, 4) == [(0, 8), (8, 16), (16, 24), (24, 33)]
    assert latitude_band_edges(3, 5) == [(0, 1), (1, 2), (2, 3)]


@pytest.mark.parametrize("bands", (0, -1))
def test_a_band_count_below_one_is_refused_by_name(bands):
    from arwen_spectral.waist import latitude_band_edges

    with pytest.raises(ValueError, match="bands must be >= 1"):
        latitude_band_edges(64, bands)
    with pytest.raises(ValueError, match="nlat must be >= 1"):
        latitude_band_edges(0, 2)


def test_rows_filled_refuses_double_fills_bad_ranges_and_gaps():
    from arwen_spectral.waist import _RowsFilled

    rows = _RowsFilled(10)
    rows.claim(0, 4)
    with pytest.raises(ValueError, match="already filled"):
        rows.claim(3, 6)
    for bad in ((5, 5), (6, 5), (-1, 2), (8, 11)):
        with pytest.raises(ValueError, match="nonempty range"):
            rows.claim(*bad)
    assert not rows.complete()
    with pytest.raises(ValueError, match="6 of 10 latitude rows"):
        rows.require_complete()
    rows.claim(4, 10)
    assert rows.complete()
    rows.require_complete()


def test_oracle_waist_of_a_zonal_harmonic_is_its_closed_form():
    """A row a cos(k lambda + phi) has waist a/2 e^{i phi} at m = k and 0
    elsewhere (m = 0: a cos phi); the drain inverts it exactly."""
    o = OracleTransform.create(21)
    lam = o.grid.lon_rad
    f = np.zeros(o.grid.shape)
    want = np.zeros((o.grid.nlat, o.truncation + 1), dtype=np.complex128)
    rng = np.random.default_rng(3)
    for j in range(o.grid.nlat):
        k = int(rng.integers(0, o.truncation + 1))
        a, phi = float(rng.normal()), float(rng.uniform(0, 2 * np.pi))
        f[j] = a * np.cos(k * lam + phi)
        want[j, k] = a * np.cos(phi) if k == 0 else 0.5 * a * np.exp(1j * phi)
    got = o.fourier_waist(f)
    assert np.max(np.abs(got - want)) < 1e-14
    assert np.max(np.abs(o.waist_to_grid(want) - f)) < 1e-13


def test_oracle_drain_ignores_the_m0_imaginary_part_and_orders_above_t_are_cut():
    o = OracleTransform.create(21)
    w = np.zeros((o.grid.nlat, o.truncation + 1), dtype=np.complex128)
    w[:, 0] = 2.0 + 5.0j
    w[:, 3] = 1.0 - 1.0j
    lam = o.grid.lon_rad
    want = 2.0 + 2.0 * (np.cos(3 * lam) + np.sin(3 * lam))
    assert np.max(np.abs(o.waist_to_grid(w) - want[None, :])) < 1e-13
    # a field with energy above T: the waist keeps m <= T only
    f = np.cos((o.truncation + 2) * lam)[None, :] * np.ones((o.grid.nlat, 1))
    assert np.max(np.abs(o.fourier_waist(f))) < 1e-14


def test_oracle_waist_routes_compose_to_the_whole_transforms():
    """Head + tail = forward, contract_to_waist + drain = inverse, the
    gradient's waists drained band by band = gradient, and the vector
    contractions of the (u, v) waist = vordiv_from_wind."""
    o = OracleTransform.create(21)
    c = red_spectrum(21, (2,), seed=4)
    f = o.inverse(c)
    assert np.array_equal(o.contract_waist(o.fourier_waist(f)), o.forward(f))
    assert np.array_equal(o.waist_to_grid(o.contract_to_waist(c)), o.inverse(c))
    east, north = o.gradient(c)
    ws = o.gradient_waists(c)
    for r0, r1 in ((0, 5), (5, 17), (17, 33)):
        e, n = o.gradient_band(ws, r0, r1)
        assert np.max(np.abs(e - east[..., r0:r1, :])) <= 1e-12 * np.max(np.abs(east))
        assert np.max(np.abs(n - north[..., r0:r1, :])) <= 1e-12 * np.max(np.abs(north))
    u, v = o.wind_from_vordiv(c[0], c[1])
    z0, d0 = o.vordiv_from_wind(u, v)
    z1, d1 = o.vordiv_from_waist(o.fourier_waist(np.stack([u, v])))
    assert np.array_equal(z0, z1) and np.array_equal(d0, d1)
    # a raw zonal spectrum wider than T+1 is cut to the retained orders
    raw = np.fft.rfft(np.stack([u, v]), axis=-1) / o.grid.nlon
    z2, d2 = o.vordiv_from_waist(raw)
    assert np.max(np.abs(z2 - z0)) < 1e-12 * np.max(np.abs(z0))
