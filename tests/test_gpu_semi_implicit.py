"""The HIP semi-implicit column solves vs the FP64 oracle (needs an MI355X).

Reference: semi_implicit.py apply (:887-919), solve_shifted (:921-959) and
_finish (:815-843).  The engine runs in float32, as the reference's device
path does with its float32 matrices; the oracle gets the same float32
matrices and complex64 spectra in float64.  Error measure per output
class: max|engine - oracle| / max|oracle|, tolerance 2e-5 (fp32 dot
products of L + 1 terms through three stages).
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import semi_implicit as osi
from oracle.sht import OracleTransform, red_spectrum

from _helpers import hip_transform

pytestmark = pytest.mark.gpu
TOL = 2e-5


def _rel(have, want):
    have = have.detach().cpu().numpy().astype(np.complex128)
    return float(np.max(np.abs(have - want)) / np.max(np.abs(want)))


@pytest.mark.parametrize("trunc,nlev,dealias", [(21, 5, 1.5), (63, 40, 1.0), (63, 64, 1.0)])
def test_vertical_solves_match_the_oracle(trunc, nlev, dealias):
    from arwen_spectral.semi_implicit import vertical_apply, vertical_solve_shifted, wave_operator

    o = OracleTransform.create(trunc, dealias_factor=dealias)
    t = hip_transform(o, "float32")
    b, c, m, modal = osi.synthetic_operator(nlev, seed=nlev)
    k2 = wave_operator(t)
    f32 = lambda a: np.asarray(a, dtype=np.float32)  # noqa: E731
    c64 = lambda a: np.asarray(a).astype(np.complex64)  # noqa: E731
    d = c64(o.project(red_spectrum(trunc, (nlev,), seed=1, scale=1e-5)))
    th = c64(o.project(red_spectrum(trunc, (nlev,), seed=2, scale=10.0)))
    ps = c64(o.project(red_spectrum(trunc, (), seed=3, scale=1e-2)))
    tw = 900.0
    inv = f32(osi.helmholtz_inverse(modal, k2, tw, 1.0))
    want = osi.solve_shifted(o, d, th, ps, f32(b), f32(c), inv, f32(k2), np.float32(tw))
    got = vertical_solve_shifted(t, d, th, ps, f32(b), f32(c), inv, f32(k2), tw)
    errs = [_rel(g, w) for g, w in zip(got, want)]
    print(f"T{trunc} L{nlev} solve_shifted rel err D {errs[0]:.2e} theta {errs[1]:.2e} "
          f"lnps {errs[2]:.2e}")
    assert max(errs) < TOL
    dt, alpha = 1200.0, 0.55
    inv = f32(osi.helmholtz_inverse(modal, k2, dt, alpha))
    want = osi.apply(o, d, th, ps, f32(b), f32(c), f32(m), inv, f32(k2), np.float32(dt), alpha)
    got = vertical_apply(t, d, th, ps, f32(b), f32(c), f32(m), inv, f32(k2), dt, alpha)
    errs = [_rel(g, w) for g, w in zip(got, want)]
    print(f"T{trunc} L{nlev} apply rel err D {errs[0]:.2e} theta {errs[1]:.2e} lnps {errs[2]:.2e}")
    assert max(errs) < TOL
    # degree zero passes through; above the triangle and Im(m = 0) are zero
    assert np.array_equal(got[0][:, 0, 0].cpu().numpy(), d[:, 0, 0].real.astype(np.complex64))
    assert not bool(got[1].cpu().numpy()[..., 0].imag.any())
    above = np.triu(np.ones((trunc + 1, trunc + 1)), 1)[None] * got[0].cpu().numpy()
    assert not above.any()


def test_vertical_solve_argument_errors():
    from arwen_spectral.semi_implicit import vertical_solve_shifted, wave_operator

    o = OracleTransform.create(21)
    t = hip_transform(o, "float32")
    L = 4
    b, c, m, modal = osi.synthetic_operator(L, seed=3)
    z = np.zeros((L, 22, 22), np.complex64)
    inv = osi.helmholtz_inverse(modal, wave_operator(t), 100.0, 1.0)
    with pytest.raises(ValueError, match="b_matrix has shape"):
        vertical_solve_shifted(t, z, z, z[0], b[:, :-1], c, inv, wave_operator(t), 100.0)
    with pytest.raises(ValueError, match="inverse has shape"):
        vertical_solve_shifted(t, z, z, z[0], b, c, inv[:-1], wave_operator(t), 100.0)
    with pytest.raises(ValueError, match="must be"):
        vertical_solve_shifted(t, z, z[:2], z[0], b, c, inv, wave_operator(t), 100.0)
    big = 65
    zb = np.zeros((big, 22, 22), np.complex64)
    with pytest.raises(ValueError, match="nlev must be in 1..64"):
        vertical_solve_shifted(t, zb, zb, z[0], np.zeros((big, big + 1)), np.zeros((big + 1, big)),
                               np.zeros((22, big, big)), wave_operator(t), 100.0)


@pytest.mark.parametrize("trunc,nlev,dealias", [(21, 5, 1.5), (63, 40, 1.0), (63, 64, 1.0)])
def test_half_map_and_tendencies_match_the_oracle(trunc, nlev, dealias):
    from arwen_spectral.semi_implicit import (vertical_half_apply, vertical_linear_tendencies,
                                              wave_operator)

    o = OracleTransform.create(trunc, dealias_factor=dealias)
    t = hip_transform(o, "float32")
    b, c, m, modal = osi.synthetic_operator(nlev, seed=nlev + 1)
    k2 = wave_operator(t)
    f32 = lambda a: np.asarray(a, dtype=np.float32)  # noqa: E731
    c64 = lambda a: np.asarray(a).astype(np.complex64)  # noqa: E731
    d = c64(o.project(red_spectrum(trunc, (nlev,), seed=4, scale=1e-5)))
    th = c64(o.project(red_spectrum(trunc, (nlev,), seed=5, scale=10.0)))
    ps = c64(o.project(red_spectrum(trunc, (), seed=6, scale=1e-2)))
    maps = [f32(x) for x in osi.symmetric_half_maps(modal, osi.x_modes(b, c), k2, 1500.0, 0.6)]
    want = osi.half_apply(o, d, th, ps, f32(b), f32(c), maps, f32(k2))
    got = vertical_half_apply(t, d, th, ps, f32(b), f32(c), *maps, f32(k2))
    errs = [_rel(g, w) for g, w in zip(got, want)]
    print(f"T{trunc} L{nlev} half_apply rel err D {errs[0]:.2e} theta {errs[1]:.2e} "
          f"lnps {errs[2]:.2e}")
    assert max(errs) < TOL
    assert np.array_equal(got[0][:, 0, 0].cpu().numpy(), d[:, 0, 0].real.astype(np.complex64))
    w = 0.7
    want = osi.linear_tendencies(o, d, th, ps, f32(b), f32(c), f32(k2), np.float32(w))
    got = vertical_linear_tendencies(t, d, th, ps, f32(b), f32(c), f32(k2), w)
    errs = [_rel(g, w_) for g, w_ in zip(got, want)]
    print(f"T{trunc} L{nlev} linear_tendencies rel err D {errs[0]:.2e} theta {errs[1]:.2e} "
          f"lnps {errs[2]:.2e}")
    assert max(errs) < TOL
