"""The HIP point sampler vs the FP64 oracle (needs an MI355X).

Reference surface: spectral/sampling.py:382-556 (sample_scalar,
sample_gradient, sample_wind, regular_latlon_scalar).  The coefficients are
complex64 (the engine's spectral dtype); the oracle evaluates the same
values in float64.  Error measure: max|engine - oracle| / max|oracle|.
Tolerances: precision "float64" 1e-12 for values and 1e-11 for gradients
(FP64 recurrence and sums; the east metric divides by cos(lat) up to 89
degrees); "float32" 2e-6 (fp32 terms, one order's sum in fp32); winds 1e-6
against the oracle, whose psi and chi are exact where the engine's, like
the reference's float32 state, are inverse_laplacian's complex64 values.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle import sampling as osamp
from oracle.sht import OracleTransform, red_spectrum

from _helpers import hip_transform

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pairs():
    out = {}
    for trunc, kw in ((21, {}), (63, {"dealias_factor": 1.0})):
        o = OracleTransform.create(trunc, **kw)
        out[trunc] = (o, hip_transform(o, "float32"))
    return out


def _points(seed, n, *, poles):
    rng = np.random.default_rng(seed)
    lat = rng.uniform(-89.0, 89.0, n)
    lon = rng.uniform(-180.0, 540.0, n)
    if poles:
        lat[:2] = [90.0, -90.0]
    return lat, lon


def _c64(c):
    return np.asarray(c).astype(np.complex64)


def _rel(have, want):
    return float(np.max(np.abs(np.asarray(have) - want)) / np.max(np.abs(want)))


@pytest.mark.parametrize("trunc", [21, 63])
@pytest.mark.parametrize("lead", [(), (3,), (2, 3)])
def test_sample_scalar_matches_the_oracle(pairs, trunc, lead):
    from arwen_spectral.sampling import sample_scalar

    o, t = pairs[trunc]
    c = _c64(red_spectrum(trunc, lead, seed=21))
    lat, lon = _points(1, 700, poles=True)
    want = osamp.sample_scalar(o, c.astype(np.complex128), lat, lon)
    got = sample_scalar(t, c, lat, lon)
    assert isinstance(got, np.ndarray) and got.dtype == np.float64 and got.shape == want.shape
    err = _rel(got, want)
    got32 = sample_scalar(t, c, lat, lon, precision="float32")
    err32 = _rel(got32, want)
    print(f"T{trunc} lead {lead} sample_scalar rel err f64 {err:.2e} f32 {err32:.2e}")
    assert err < 1e-12 and err32 < 2e-6
    # point shape kept, chunking changes no bit, device results on request
    lat2, lon2 = lat[:600].reshape(20, 30), lon[:600].reshape(20, 30)
    a = sample_scalar(t, c, lat2, lon2)
    b = sample_scalar(t, c, lat2, lon2, chunk_points=77)
    assert a.shape == (*lead, 20, 30) and np.array_equal(a, b)
    d = sample_scalar(t, c, lat2, lon2, as_device=True)
    assert d.is_cuda and np.array_equal(d.cpu().numpy(), a)


@pytest.mark.parametrize("trunc", [21, 63])
def test_sample_gradient_and_wind_match_the_oracle(pairs, trunc):
    from arwen_spectral.sampling import sample_gradient, sample_wind

    o, t = pairs[trunc]
    c = _c64(red_spectrum(trunc, (3,), seed=22))
    lat, lon = _points(2, 500, poles=False)
    we, wn = osamp.sample_gradient(o, c.astype(np.complex128), lat, lon)
    ge, gn = sample_gradient(t, c, lat, lon)
    err = max(_rel(ge, we), _rel(gn, wn))
    ge32, gn32 = sample_gradient(t, c, lat, lon, precision="float32")
    err32 = max(_rel(ge32, we), _rel(gn32, wn))
    print(f"T{trunc} sample_gradient rel err f64 {err:.2e} f32 {err32:.2e}")
    assert err < 1e-11 and err32 < 2e-6

    z = _c64(red_spectrum(trunc, (2,), seed=23, scale=1e-5))
    d = _c64(red_spectrum(trunc, (2,), seed=24, scale=2e-6))
    wu, wv = osamp.sample_wind(o, z.astype(np.complex128), d.astype(np.complex128), lat, lon)
    u, v = sample_wind(t, z, d, lat, lon)
    werr = max(_rel(u, wu), _rel(v, wv))
    print(f"T{trunc} sample_wind rel err {werr:.2e}")
    assert werr < 1e-6
    # one launch for the wind = the reference's route through inverse_laplacian
    pe, pn = sample_gradient(t, t.inverse_laplacian(z), lat, lon)
    ce, cn = sample_gradient(t, t.inverse_laplacian(d), lat, lon)
    assert np.array_equal(u, -pn + ce) and np.array_equal(v, pe + cn)


def test_regular_grid_and_argument_errors(pairs):
    from arwen_spectral import _native
    from arwen_spectral.sampling import (regular_latlon_scalar, sample_gradient, sample_scalar,
                                         sample_wind)

    o, t = pairs[21]
    c = _c64(red_spectrum(21, (2,), seed=25))
    lat, lon, f = regular_latlon_scalar(t, c, nlat=9, nlon=16, include_poles=True)
    lon2, lat2 = np.meshgrid(lon, lat)
    assert f.shape == (2, 9, 16)
    assert _rel(f, osamp.sample_scalar(o, c.astype(np.complex128), lat2, lon2)) < 1e-12
    with pytest.raises(ValueError, match="undefined at the exact poles"):
        sample_gradient(t, c, [90.0], [0.0])
    with pytest.raises(ValueError, match="latitude must lie"):
        sample_scalar(t, c, [-90.5], [0.0])
    with pytest.raises(ValueError, match="shape mismatch"):
        sample_wind(t, c, c[:1], [10.0], [0.0])
    with pytest.raises(ValueError, match="trailing shape"):
        sample_scalar(t, np.zeros((3, 3), np.complex64), [10.0], [0.0])
    empty = sample_scalar(t, c, np.zeros((0,)), np.zeros((0,)))
    assert empty.shape == (2, 0)
    # the C ABI refuses a bad kind, a bad precision and a short workspace
    import torch

    lib = _native.load()
    dev = torch.device("cuda")
    cc = torch.from_numpy(c).to(dev)
    pts = torch.zeros(4, dtype=torch.float64, device=dev)
    out = torch.empty((2, 4), dtype=torch.float64, device=dev)
    need = int(lib.arwen_sht_sample_workspace_bytes(t.plan_handle, 2, 0, 4))
    assert need == (22 * 23 // 2) * 8 * 16  # T21: one order-range part, no partial sums
    work = torch.empty(need, dtype=torch.uint8, device=dev)
    vp = lambda x: ctypes.c_void_p(x.data_ptr())  # noqa: E731
    for kind, prec, nbytes, msg in ((3, 0, need, b"kind"), (0, 2, need, b"precision"),
                                    (0, 0, need - 1, b"workspace")):
        rc = lib.arwen_sht_sample(t.plan_handle, vp(cc), None, 2, vp(pts), vp(pts), vp(pts), 4,
                                  kind, prec, vp(work), nbytes, vp(out), None, t.backend.stream())
        assert rc == _native.EINVAL and msg in lib.arwen_sht_last_error()


def test_t383_sample_on_a_subset(pairs):
    """The longest recurrences of the bench truncation, both precisions."""
    from arwen_spectral import SphericalHarmonicTransform
    from arwen_spectral.sampling import sample_scalar

    o = OracleTransform.create(383, dealias_factor=1.0)
    t = SphericalHarmonicTransform.create(383, dealias_factor=1.0, backend="hip")
    c = _c64(red_spectrum(383, (2,), seed=26))
    lat, lon = _points(3, 300, poles=True)
    want = osamp.sample_scalar(o, c.astype(np.complex128), lat, lon)
    err = _rel(sample_scalar(t, c, lat, lon), want)
    err32 = _rel(sample_scalar(t, c, lat, lon, precision="float32"), want)
    print(f"T383 sample_scalar rel err f64 {err:.2e} f32 {err32:.2e}")
    assert err < 1e-12 and err32 < 2e-6
