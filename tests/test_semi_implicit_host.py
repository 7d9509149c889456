"""The oracle's semi-implicit column solves pinned by the linear systems they
solve (reference semi_implicit.py:887-959): with L y = (k^2 B x, C D),
solve_shifted returns y' with (I - tau w L) y' = y, and apply returns the
off-centred Crank-Nicolson step (I - alpha dt L) y' = (I + (1-alpha) dt L) y,
for every retained column of degree n >= 1; degree zero passes through and
the outputs are projected.  The Helmholtz inverses here are numpy's
``linalg.inv`` of (I + s k_n^2 M), independent of the oracle's modal route.
"""
from __future__ import annotations

import numpy as np

from oracle import semi_implicit as osi
from oracle.sht import OracleTransform, red_spectrum


def _case(T=10, L=6, seed=1):
    o = OracleTransform.create(T)
    b, c, m, modal = osi.synthetic_operator(L, seed=seed)
    n = np.arange(T + 1, dtype=np.float64)
    k2 = n * (n + 1.0) / o.grid.radius_m ** 2
    d = o.project(red_spectrum(T, (L,), seed=seed + 1, scale=1e-5))
    th = o.project(red_spectrum(T, (L,), seed=seed + 2, scale=10.0))
    ps = o.project(red_spectrum(T, (), seed=seed + 3, scale=1e-2))
    return o, b, c, m, modal, k2, d, th, ps


def _inverses(m, k2, s):
    eye = np.eye(m.shape[0])
    return np.stack([np.linalg.inv(eye + s * kk * m) for kk in k2])


def _residual(lhs, rhs):
    tri = np.tri(lhs.shape[-1], dtype=bool)
    tri[0, 0] = False  # degree zero is pinned, not solved
    return float(np.max(np.abs((lhs - rhs)[..., tri])) / np.max(np.abs(rhs[..., tri])))


def test_modal_helmholtz_inverse_is_the_matrix_inverse():
    o, b, c, m, modal, k2, *_ = _case()
    assert np.allclose(-b @ c, m, rtol=0, atol=1e-9 * np.max(np.abs(m)))
    inv = osi.helmholtz_inverse(modal, k2, 600.0, 0.6)
    want = _inverses(m, k2, (0.6 * 600.0) ** 2)
    assert np.max(np.abs(inv - want)) < 1e-12 * np.max(np.abs(want))
    assert np.array_equal(inv[0], np.eye(m.shape[0]))


def test_solve_shifted_solves_the_implicit_stage():
    o, b, c, m, modal, k2, d, th, ps = _case()
    tw = 900.0
    d1, th1, ps1 = osi.solve_shifted(o, d, th, ps, b, c, _inverses(m, k2, tw * tw), k2, tw)
    x = np.concatenate([th, ps[None]])
    x1 = np.concatenate([th1, ps1[None]])
    k2c = k2[None, :, None]
    # d' - tw k^2 B x' = D ; x' - tw C d' = x
    assert _residual(d1 - tw * k2c * np.einsum("ij,jnm->inm", b, x1), d) < 1e-12
    assert _residual(x1 - tw * np.einsum("ij,jnm->inm", c, d1), x) < 1e-12
    assert np.array_equal(d1[:, 0, 0], d[:, 0, 0].real) and np.array_equal(x1[:, 0, 0], x[:, 0, 0].real)
    assert not np.any(np.triu(np.ones((11, 11)), 1)[None] * d1)
    assert not np.any(d1[..., 0].imag) and not np.any(ps1[:, 0].imag)


def test_apply_is_the_off_centred_crank_nicolson_step():
    o, b, c, m, modal, k2, d, th, ps = _case(seed=4)
    dt, alpha = 1200.0, 0.55
    inv = osi.helmholtz_inverse(modal, k2, dt, alpha)
    d1, th1, ps1 = osi.apply(o, d, th, ps, b, c, m, inv, k2, dt, alpha)
    x = np.concatenate([th, ps[None]])
    x1 = np.concatenate([th1, ps1[None]])
    k2c = k2[None, :, None]
    bx = lambda v: k2c * np.einsum("ij,jnm->inm", b, v)  # noqa: E731
    cd = lambda v: np.einsum("ij,jnm->inm", c, v)  # noqa: E731
    assert _residual(d1 - alpha * dt * bx(x1), d + (1 - alpha) * dt * bx(x)) < 1e-12
    assert _residual(x1 - alpha * dt * cd(d1), x + (1 - alpha) * dt * cd(d)) < 1e-12


def test_half_map_applied_twice_is_the_crank_nicolson_step():
    """H^2 = CN (semi_implicit.py:845-855): the symmetric split composes to
    the off-centred Crank-Nicolson step exactly."""
    o, b, c, m, modal, k2, d, th, ps = _case(seed=7)
    dt, alpha = 1500.0, 0.6
    maps = osi.symmetric_half_maps(modal, osi.x_modes(b, c), k2, dt, alpha)
    once = osi.half_apply(o, d, th, ps, b, c, maps, k2)
    twice = osi.half_apply(o, *once, b, c, maps, k2)
    full = osi.apply(o, d, th, ps, b, c, m, osi.helmholtz_inverse(modal, k2, dt, alpha), k2, dt,
                     alpha)
    for got, want in zip(twice, full):
        assert np.max(np.abs(got - want)) < 1e-11 * np.max(np.abs(want))


def test_linear_tendencies_are_the_first_order_term_of_the_step():
    """(apply_dt(y) - y) / dt -> L y as dt -> 0, with error O(dt)."""
    o, b, c, m, modal, k2, d, th, ps = _case(seed=8)
    lt = osi.linear_tendencies(o, d, th, ps, b, c, k2, 1.0)
    errs = []
    for dt in (1e-2, 1e-3):
        step = osi.apply(o, d, th, ps, b, c, m, osi.helmholtz_inverse(modal, k2, dt, 0.5), k2,
                         dt, 0.5)
        diffs = [(s - y) / dt for s, y in zip(step, (o.project(d), o.project(th), o.project(ps)))]
        tri = np.tri(11, dtype=bool)
        tri[0, 0] = False
        errs.append(max(float(np.max(np.abs((g - w)[..., tri])) / np.max(np.abs(w[..., tri])))
                        for g, w in zip(diffs, lt)))
    assert errs[0] < 1e-3 and errs[1] < errs[0] / 5
