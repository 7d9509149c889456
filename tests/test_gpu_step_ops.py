"""The dycore's per-step operations through the C ABI (needs an MI355X).

  * inverse with both gradients from one pack (``arwen_sht_synthesis_gradient``)
    against the FP64 oracle's inverse and gradient (reference
    transform.py:1742-1768);
  * ln ps about ln(p_ref): ``inverse_about`` with the (0, 0) offset fused
    into the pack is bit-identical to ``inverse(lnps_anomaly_coefficients(c))``
    (reference lnps_reference.py:32-49, dynamics.py:964-980), and the exp
    epilogue matches exp of the inverse;
  * theta analysed about whole-kelvin offsets: ``forward_about`` is
    bit-identical to the reference's float32 sequence
    ``forward(g - c) ; [..., 0, 0] += c sqrt(4 pi)`` (semilag/step.py:128-163,
    610-621), within the FP32 tolerance of the oracle's ``forward(g)``, and a
    constant plane at the offset analyses to its (0, 0) coefficient alone
    where the plain analysis leaks into m = 0;
  * the same options inside a TransformBatch;
  * repeated single operations with fresh buffers write the right results
    and allocate no device memory after the first call of a shape.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle.sht import OracleTransform, red_spectrum

from _helpers import hip_transform, rel_err

pytestmark = pytest.mark.gpu

TOL = {"float32": 1e-6, "bf16x6": 1e-6}


@pytest.fixture(scope="module")
def oracles():
    return {
        21: OracleTransform.create(21),                       # 33 x 66: Stockham FFT path
        63: OracleTransform.create(63, dealias_factor=1.0),   # 64 x 128: wave-resident FFT path
        127: OracleTransform.create(127),                     # 192 x 384: wave FFT, 6 passes
    }


def _dev(a, dtype):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a).astype(dtype)).to(
        torch.device("cuda", torch.cuda.current_device()))


def _np(t):
    return t.detach().cpu().numpy()


def _same(a, b) -> bool:
    """Bit-identical arrays; prints the largest difference when not."""
    a, b = _np(a), _np(b)
    if a.shape == b.shape and np.array_equal(a, b):
        return True
    print(f"differs: shapes {a.shape} {b.shape}, max |a - b| {np.max(np.abs(a - b)):.3e}")
    return False


@pytest.mark.parametrize("precision", ["float32", "bf16x6"])
@pytest.mark.parametrize("T", [21, 63, 127])
def test_inverse_with_gradient_matches_oracle(oracles, T, precision):
    o = oracles[T]
    tr = hip_transform(o, precision)
    c = o.project(red_spectrum(T, (3,), seed=7, scale=30.0, mean=280.0))
    grid, east, north = tr.inverse_with_gradient(_dev(c, np.complex64))
    want_e, want_n = o.gradient(c)
    e_g = rel_err(grid, o.inverse(c))
    e_e = rel_err(east, want_e)
    e_n = rel_err(north, want_n)
    print(f"T{T} {precision}: grid {e_g:.2e} east {e_e:.2e} north {e_n:.2e}")
    assert e_g <= TOL[precision]
    assert e_e <= TOL[precision]
    assert e_n <= TOL[precision]
    # the same as the separate operations (the north derivative, and the
    # grid, come from the same tables and pack: bit-identical)
    sep_e, sep_n = tr.gradient(_dev(c, np.complex64))
    assert _same(grid, tr.inverse(_dev(c, np.complex64)))
    assert _same(north, sep_n)
    # the east derivative takes i m before the contraction in the separate
    # operation and after it here: two float32 roundings of the same sum
    e_sep = rel_err(east, _np(sep_e).astype(np.float64))
    assert e_sep <= 5e-7


@pytest.mark.parametrize("precision", ["float32", "bf16x6"])
@pytest.mark.parametrize("T", [21, 63, 127])
def test_lnps_anomaly_inverse_and_exp(oracles, T, precision):
    import torch
    from arwen_spectral import LNPS_REFERENCE, lnps_anomaly_coefficients
    from arwen_spectral.offsets import lnps_offset

    o = oracles[T]
    tr = hip_transform(o, precision)
    lnps = o.project(red_spectrum(T, (2,), seed=3, scale=0.02))
    lnps[..., 0, 0] = math.log(98_500.0) * math.sqrt(4.0 * math.pi)
    c = _dev(lnps, np.complex64)
    # the reference's sequence: the constant taken off (0, 0) in complex64, then inverse
    want = tr.inverse(lnps_anomaly_coefficients(tr, c))
    have = tr.inverse_about(c, lnps_offset(-1.0))
    assert _same(have, want)
    # and against the oracle on the same float32 coefficients
    c32 = _np(lnps_anomaly_coefficients(tr, c)).astype(np.complex128)
    e = rel_err(have, o.inverse(c32))
    print(f"T{T} {precision}: anomaly {e:.2e}")
    assert e <= TOL[precision]
    # exp epilogue: ps = exp(ln ps)
    ps = tr.inverse_about(c, None, exp=True)
    err = rel_err(ps, torch.exp(tr.inverse(c)).double().cpu().numpy())
    assert err <= 4e-7, err
    # the (0, 0) anomaly is the ln ps mean less the fixed ln(p_ref)
    mean = float(np.real(_np(lnps_anomaly_coefficients(tr, c))[0, 0, 0])) / math.sqrt(4 * math.pi)
    assert abs(mean - (math.log(98_500.0) - LNPS_REFERENCE)) < 1e-5


@pytest.mark.parametrize("precision", ["float32", "bf16x6"])
@pytest.mark.parametrize("T", [21, 63, 127])
def test_theta_forward_about_offsets(oracles, T, precision):
    import torch
    from arwen_spectral import theta_analysis_offsets

    o = oracles[T]
    tr = hip_transform(o, precision)
    column = np.array([300.4, 318.7, 512.2, 1799.6])
    grid_off, coef_off = theta_analysis_offsets(tr, column)
    assert _np(grid_off).tolist() == [300.0, 319.0, 512.0, 1800.0]
    anomaly = o.inverse(o.project(red_spectrum(T, (4,), seed=11, scale=2.0)))
    theta = (column[:, None, None] + anomaly).astype(np.float32)
    g = _dev(theta, np.float32)
    have = tr.forward_about(g, grid_off, coef_off)
    # the reference's float32 sequence, op for op
    want = tr.forward(g - grid_off[:, None, None])
    want[:, 0, 0] += coef_off
    assert _same(have, want)
    e = rel_err(have, o.forward(theta.astype(np.float64)))
    print(f"T{T} {precision}: theta about offsets {e:.2e}")
    assert e <= TOL[precision]
    # a constant plane at the offset: (0, 0) alone, no m = 0 leak
    plane = torch.full((1, *tr.grid.shape), 300.0, dtype=torch.float32, device=g.device)
    about = _np(tr.forward_about(plane, 300.0, 300.0 * math.sqrt(4 * math.pi)))[0]
    plain = _np(tr.forward(plane))[0]
    assert about[0, 0] == np.complex64(300.0 * math.sqrt(4 * math.pi))
    rest = about.copy()
    rest[0, 0] = 0
    assert not np.any(rest)
    leak_plain = float(np.max(np.abs(plain[1:, 0])))
    print(f"T{T} {precision}: plain m = 0 leak of a 300 K plane {leak_plain:.2e}")


def test_offsets_and_gradient_inside_a_batch(oracles):
    import torch
    from arwen_spectral import TransformBatch, theta_analysis_offsets
    from arwen_spectral.offsets import lnps_offset

    o = oracles[127]
    tr = hip_transform(o, "bf16x6")
    T = o.grid.truncation
    dev = torch.device("cuda", torch.cuda.current_device())
    lnps = o.project(red_spectrum(T, (1,), seed=5, scale=0.02))
    lnps[..., 0, 0] += math.log(99_000.0) * math.sqrt(4.0 * math.pi)
    temp = o.project(red_spectrum(T, (3,), seed=6, scale=20.0, mean=250.0))
    c_lnps = _dev(lnps, np.complex64)
    c_temp = _dev(temp, np.complex64)
    column = np.array([290.2, 330.8, 700.1])
    theta = (column[:, None, None] + o.inverse(o.project(red_spectrum(T, (3,), seed=8)))) \
        .astype(np.float32)
    g_theta = _dev(theta, np.float32)
    grid_off, coef_off = theta_analysis_offsets(tr, column)
    anom_off = torch.full((1,), lnps_offset(-1.0), dtype=torch.complex64, device=dev)
    g = lambda n: torch.empty((n, *tr.grid.shape), dtype=torch.float32, device=dev)
    s = lambda n: torch.empty((n, *tr.spectral_shape), dtype=torch.complex64, device=dev)
    out = dict(logps=g(1), dlx=g(1), dly=g(1), ps=g(1), anom=g(1), t=g(3), dtx=g(3), dty=g(3),
               theta=s(3))
    batch = TransformBatch([
        ("inverse_gradient", tr, c_lnps, (out["logps"], out["dlx"], out["dly"])),
        ("inverse", tr, c_lnps, out["ps"], {"exp": True}),
        ("inverse", tr, c_lnps, out["anom"], {"coeff_offset": anom_off}),
        ("inverse_gradient", tr, c_temp, (out["t"], out["dtx"], out["dty"])),
        ("forward", tr, g_theta, out["theta"], {"grid_offset": grid_off, "coeff_offset": coef_off}),
    ])
    for _ in range(2):
        batch.run()
        torch.cuda.synchronize()
        assert _same(out["logps"], tr.inverse(c_lnps))
        assert _same(out["ps"], tr.inverse_about(c_lnps, None, exp=True))
        assert _same(out["anom"], tr.inverse_about(c_lnps, lnps_offset(-1.0)))
        assert _same(out["theta"], tr.forward_about(g_theta, grid_off, coef_off))
        for k, (grid, east, north) in (("lnps", (out["logps"], out["dlx"], out["dly"])),
                                       ("temp", (out["t"], out["dtx"], out["dty"]))):
            c = lnps if k == "lnps" else temp
            we, wn = o.gradient(c)
            err = rel_err(grid, o.inverse(c))
            assert err <= 1e-6, err
            err = rel_err(east, we)
            assert err <= 1e-6, err
            err = rel_err(north, wn)
            assert err <= 1e-6, err
    with pytest.raises(ValueError):
        TransformBatch([("forward", tr, g_theta, out["theta"], {"exp": True})])
    with pytest.raises(ValueError):
        TransformBatch([("forward", tr, g_theta, out["theta"], {"grid_offset": grid_off[:2]})])


def test_single_ops_rebind_buffers_without_allocating(oracles):
    import ctypes

    import torch
    from arwen_spectral import _native

    o = oracles[63]
    tr = hip_transform(o, "float32")
    lib = _native.load()
    stream = tr.backend.stream()
    T = o.grid.truncation
    dev = torch.device("cuda", torch.cuda.current_device())
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    coeffs = [o.project(red_spectrum(T, (5,), seed=20 + k, scale=1.0)) for k in range(4)]
    ins = [_dev(c, np.complex64) for c in coeffs]
    outs = [torch.empty((5, *tr.grid.shape), dtype=torch.float32, device=dev) for _ in range(4)]
    east = [torch.empty_like(outs[0]) for _ in range(4)]
    north = [torch.empty_like(outs[0]) for _ in range(4)]
    # the first call of each key sets up the cached batch
    _native.check(lib.arwen_sht_synthesis(tr.plan_handle, vp(ins[0]), vp(outs[0]), 5, stream))
    _native.check(lib.arwen_sht_gradient(tr.plan_handle, vp(ins[0]), vp(east[0]), vp(north[0]),
                                         5, stream))
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for k in range(1, 4):
        _native.check(lib.arwen_sht_synthesis(tr.plan_handle, vp(ins[k]), vp(outs[k]), 5, stream))
        _native.check(lib.arwen_sht_gradient(tr.plan_handle, vp(ins[k]), vp(east[k]),
                                             vp(north[k]), 5, stream))
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
    for k in range(4):
        we, wn = o.gradient(coeffs[k])
        err = rel_err(outs[k], o.inverse(coeffs[k]))
        assert err <= 1e-6, err
        err = rel_err(east[k], we)
        assert err <= 1e-6, err
        err = rel_err(north[k], wn)
        assert err <= 1e-6, err
