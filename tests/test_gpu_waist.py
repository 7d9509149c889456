"""The Fourier waist API on the HIP engine vs the FP64 oracle (needs an MI355X).

Covers the reference's band pipeline surface (spectral/transform.py:68-629,
1276-1472, 1629-1796; vector.py:75-165): fill / close / contract, contract
/ drain, banded drains, the gradient's waists, the vector analysis from a
waist, the waist's lifetime and the errors.  Error measure as in the parity
tests, max|engine - oracle| / max|oracle|; the waist maps themselves are
float32 FFTs (tolerance 1e-6), the contractions keep the parity tolerances
(float32 / bf16x6 2e-6).  Band fills and band drains are compared with the
whole-field call bit for bit: every row is transformed on its own.

Grids: T21 on 33 x 66 (odd nlat, so the equator row folds onto itself; nlon
is not a multiple of 64, so the waist maps take the direct-sum kernels) and
T63 on 64 x 128 (the wave-resident FFT kernels).
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle.sht import OracleTransform, red_spectrum

from _helpers import hip_transform, rel_err

pytestmark = pytest.mark.gpu

FFT_TOL = 1e-6
TOL = {"float32": 2e-6, "bf16x6": 2e-6}


@pytest.fixture(scope="module")
def oracles():
    return {
        21: OracleTransform.create(21),
        63: OracleTransform.create(63, dealias_factor=1.0),
    }


@pytest.fixture(scope="module")
def engines(oracles):
    return {(t, p): hip_transform(o, p) for t, o in oracles.items() for p in TOL}


def _field(o, lead, seed):
    return o.inverse(red_spectrum(o.truncation, lead, seed=seed))


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("trunc", [21, 63])
def test_fourier_waist_matches_the_oracle_and_band_fills_are_bit_identical(oracles, engines,
                                                                           trunc):
    import torch

    o = oracles[trunc]
    t = engines[(trunc, "float32")]
    f = _field(o, (2, 3), seed=11).astype(np.float32)
    whole = t.fourier_waist(f)
    assert whole.shape == (2, 3, o.grid.nlat, trunc + 1)
    assert whole.values.is_contiguous() and whole.values.dtype == torch.complex64
    err = rel_err(whole.values, o.fourier_waist(f.astype(np.float64)))
    print(f"T{trunc} fourier_waist rel err {err:.2e}")
    assert err < FFT_TOL
    for bands in (2, 3, 7):
        w = t.fourier_waist(f, bands=bands)
        assert w.bands == bands and torch.equal(w.values, whole.values)
        # an explicit band pipeline over the transform's own schedule
        p = t.open_waist((2, 3), bands=bands)
        for r0, r1 in reversed(p.edges):  # any order
            p.fill_band(r0, r1, f[..., r0:r1, :])
        assert torch.equal(p.close().values, whole.values), f"{bands} bands"


@pytest.mark.parametrize("trunc", [21, 63])
@pytest.mark.parametrize("precision", list(TOL))
def test_waist_head_and_tail_reproduce_the_analysis(oracles, engines, trunc, precision):
    o = oracles[trunc]
    t = engines[(trunc, precision)]
    f = _field(o, (4,), seed=12)
    want = o.forward(f)
    p = t.open_waist((4,), bands=3)
    for r0, r1 in p.edges:
        p.fill_band(r0, r1, f[..., r0:r1, :].astype(np.float32))
    got = t.contract_waist(p.close())
    err = rel_err(got, want)
    print(f"T{trunc} {precision} contract_waist rel err {err:.2e}")
    assert err < TOL[precision]
    assert float(np.max(np.abs(_np(got)[..., :, 0].imag))) == 0.0
    with pytest.raises(RuntimeError, match="consumed"):
        p.values
    # contract_values on a leading slice of a waist's buffer
    w = t.fourier_waist(f.astype(np.float32))
    part = t.contract_values(w.values[1:3])
    assert rel_err(part, want[1:3]) < TOL[precision]
    assert rel_err(t.forward(f.astype(np.float32)), want) < TOL[precision]


@pytest.mark.parametrize("trunc", [21, 63])
@pytest.mark.parametrize("precision", list(TOL))
def test_contract_to_waist_and_drains_reproduce_the_synthesis(oracles, engines, trunc,
                                                              precision):
    import torch

    o = oracles[trunc]
    t = engines[(trunc, precision)]
    c = red_spectrum(trunc, (2, 2), seed=13)
    w = t.contract_to_waist(c, bands=4)
    assert w.bands == 4
    err_w = rel_err(w.values, o.contract_to_waist(c))
    assert err_w < TOL[precision], err_w
    pieces = [t.waist_band_to_grid(w, r0, r1) for r0, r1 in w.edges]
    again = [t.waist_band_to_grid(w, r0, r1) for r0, r1 in reversed(w.edges)][::-1]
    whole = t.waist_to_grid(w)
    assert torch.equal(torch.cat(pieces, dim=-2), whole)
    assert torch.equal(torch.cat(again, dim=-2), whole)
    err = rel_err(whole, o.inverse(c))
    print(f"T{trunc} {precision} waist_to_grid rel err {err:.2e}")
    assert err < TOL[precision]
    with pytest.raises(RuntimeError, match="consumed"):
        t.waist_to_grid(w)
    assert w.shape == (2, 2, o.grid.nlat, trunc + 1) and w.bands == 4
    dw = t.contract_to_waist(c, t._derivative_basis)
    assert rel_err(dw.values, o.contract_to_waist(c, "derivative")) < TOL[precision]
    assert rel_err(t.waist_to_grid(dw), o.inverse_meridional_derivative(c)) < TOL[precision]


@pytest.mark.parametrize("trunc", [21, 63])
def test_gradient_waists_drain_band_by_band_to_the_gradient(oracles, engines, trunc):
    o = oracles[trunc]
    t = engines[(trunc, "bf16x6")]
    c = red_spectrum(trunc, (3,), seed=14)
    east, north = o.gradient(c)
    ws = t.gradient_waists(c, bands=3)
    assert all(w.bands == 3 for w in ws)
    got_e, got_n = [], []
    for r0, r1 in ws[0].edges:
        e, n = t.gradient_band(ws, r0, r1)
        got_e.append(_np(e))
        got_n.append(_np(n))
    err_e = rel_err(np.concatenate(got_e, axis=-2), east)
    err_n = rel_err(np.concatenate(got_n, axis=-2), north)
    print(f"T{trunc} gradient_band rel err east {err_e:.2e} north {err_n:.2e}")
    assert err_e < TOL["bf16x6"] and err_n < TOL["bf16x6"]


@pytest.mark.parametrize("trunc", [21, 63])
@pytest.mark.parametrize("precision", list(TOL))
def test_vordiv_from_a_banded_wind_waist(oracles, engines, trunc, precision):
    from arwen_spectral import VorticityDivergenceOperator

    o = oracles[trunc]
    t = engines[(trunc, precision)]
    op = VorticityDivergenceOperator(t)
    cz = red_spectrum(trunc, (2,), seed=15)
    cd = red_spectrum(trunc, (2,), seed=16, scale=0.3)
    u, v = o.wind_from_vordiv(cz, cd)
    pair = np.stack([u, v]).astype(np.float32)
    zeta, div = o.vordiv_from_wind(pair[0].astype(np.float64), pair[1].astype(np.float64))
    scale = max(float(np.max(np.abs(zeta))), float(np.max(np.abs(div))))

    def pair_err(z, d):  # (zeta, D) is one vector field: errors against its joint scale
        return max(float(np.max(np.abs(_np(z) - zeta))), float(np.max(np.abs(_np(d) - div)))) / scale

    p = t.open_waist((2, 2), bands=5)
    for r0, r1 in p.edges:
        p.fill_band(r0, r1, pair[..., r0:r1, :])
    got_z, got_d = op.vordiv_from_wind_waist(p.close())
    err = pair_err(got_z, got_d)
    print(f"T{trunc} {precision} vordiv_from_wind_waist rel err {err:.2e}")
    assert err < TOL[precision]
    # the waist is read, not consumed; the divergence alone (two plane
    # contractions per level, the pair's own products) agrees with the pair's
    d_only = op.divergence_from_fourier(p)
    assert pair_err(got_z, d_only) < TOL[precision]
    same = np.array_equal(_np(d_only), _np(got_d))
    gap = float(np.max(np.abs(_np(d_only) - _np(got_d)))) / scale
    print(f"T{trunc} {precision} divergence_from_fourier vs the pair: identical {same}, "
          f"gap {gap:.2e}")
    assert gap <= 1e-7
    # a raw zonal spectrum wider than T+1 is cut to the retained orders
    import torch

    raw = torch.fft.rfft(torch.from_numpy(pair).to(p.values.device), dim=-1) / o.grid.nlon
    assert pair_err(*op._vordiv_from_fourier(raw)) < TOL[precision]
    w = op._wind_fourier(pair, bands=2)
    assert w.shape == (2, 2, o.grid.nlat, trunc + 1) and w.bands == 2


def test_waist_errors_are_refused_by_name(oracles, engines):
    from arwen_spectral import _native

    o = oracles[21]
    t = engines[(21, "float32")]
    nlat, nlon = o.grid.shape
    f = np.zeros((2, nlat, nlon), dtype=np.float32)
    p = t.open_waist((2,), bands=2)
    p.fill_band(0, 10, f[:, :10])
    with pytest.raises(ValueError, match="already filled"):
        p.fill_band(5, 12, f[:, 5:12])
    with pytest.raises(ValueError, match="are not the waist's"):
        p.fill_band(10, 12, f[:, 10:13])
    with pytest.raises(ValueError, match="were never filled"):
        t.contract_waist(p)
    full = t.fourier_waist(f)
    with pytest.raises(RuntimeError, match="not opened for band fills"):
        full.fill_band(0, 1, f[:, :1])
    with pytest.raises(ValueError, match="contracts with the plan's analysis"):
        t.contract_waist(full, t._basis)
    with pytest.raises(ValueError, match="basis, derivative"):
        t.contract_to_waist(np.zeros((22, 22), np.complex64), t._analysis)
    with pytest.raises(ValueError, match="nonempty range"):
        t.waist_band_to_grid(full, 5, 5)
    with pytest.raises(ValueError, match="trailing shape"):
        t.contract_values(np.zeros((2, nlat, 23), np.complex64))
    scratch = full.drain_scratch(3)
    assert tuple(scratch.shape) == (2, 3, nlon // 2 + 1) and not bool(scratch.abs().sum())
    # the C ABI refuses an unknown kind and rows outside the waist
    lib = _native.load()
    c = t.zeros(1)
    out = t._empty_waist((1,))
    rc = lib.arwen_sht_synthesis_to_waist(t.plan_handle, ctypes.c_void_p(c.data_ptr()),
                                          ctypes.c_void_p(out.data_ptr()), 1, 7,
                                          t.backend.stream())
    assert rc == _native.EINVAL and b"kind" in lib.arwen_sht_last_error()
    rows = t._empty_grid((1,))
    rc = lib.arwen_sht_rows_from_waist(t.plan_handle, ctypes.c_void_p(out.data_ptr()),
                                       ctypes.c_void_p(rows.data_ptr()), 1, 4, nlat - 3,
                                       t.backend.stream())
    assert rc == _native.EINVAL and b"outside the waist" in lib.arwen_sht_last_error()
