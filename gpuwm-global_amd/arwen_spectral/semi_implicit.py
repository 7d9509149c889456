"""The semi-implicit vertical-mode solves on spectral columns, on MI355X.

The reference's ``VerticalModeSemiImplicit`` (src/arwen_global/
semi_implicit.py) builds its vertical operator on the host in FP64 -- the
matrices B (L x L+1), C (L+1 x L), M = -B C and the per-degree Helmholtz
inverses (I + s k_n^2 M)^-1 -- and casts them to the state's float dtype;
each step then runs the same few contractions over the (L, T+1, T+1)
spectra.  These functions are those device contractions: the caller hands
over the matrices its operator built (numpy or device arrays), the columns
are solved by one HIP kernel (``arwen_sht_vertical_solve``), and degree zero
is pinned and the outputs projected as the reference's ``_finish`` does.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .transform import SphericalHarmonicTransform, _vp


def wave_operator(transform: SphericalHarmonicTransform) -> np.ndarray:
    """k_n^2 = n (n + 1) / a^2 per total degree (reference semi_implicit.py:792-794)."""
    n = np.arange(transform.truncation + 1, dtype=np.float64)
    return n * (n + 1.0) / (transform.grid.radius_m ** 2)


def _matrix(t, a, shape, name):
    m = t.backend.asarray(a, dtype=t.backend.float_dtype).contiguous()
    if tuple(m.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(m.shape)}, expected {tuple(shape)}")
    return m


def _columns(t: SphericalHarmonicTransform, divergence, theta, log_surface_pressure, b_matrix,
             c_matrix):
    d = t._spectral_tensor(divergence)
    th = t._spectral_tensor(theta)
    ps = t._spectral_tensor(log_surface_pressure)
    if d.dim() != 3 or th.shape != d.shape or ps.dim() != 2:
        raise ValueError(
            "divergence and theta must be (nlev, T+1, T+1) and ln ps (T+1, T+1), got "
            f"{tuple(d.shape)}, {tuple(th.shape)}, {tuple(ps.shape)}")
    L = int(d.shape[0])
    b = _matrix(t, b_matrix, (L, L + 1), "b_matrix")
    c = _matrix(t, c_matrix, (L + 1, L), "c_matrix")
    outs = (t._empty_spectral((L,)), t._empty_spectral((L,)), t._empty_spectral(()))
    return d, th, ps, L, b, c, outs


def _solve(t: SphericalHarmonicTransform, divergence, theta, log_surface_pressure, b_matrix,
           c_matrix, m_matrix, inverse, k2, s_bx, s_md, s_c, s_cd, s_cn):
    d, th, ps, L, b, c, (d_out, th_out, ps_out) = _columns(
        t, divergence, theta, log_surface_pressure, b_matrix, c_matrix)
    T1 = t.truncation + 1
    m = _matrix(t, m_matrix, (L, L), "structure") if m_matrix is not None else None
    h = _matrix(t, inverse, (T1, L, L), "inverse")
    k = _matrix(t, k2, (T1,), "k2")
    _native.check(t._lib.arwen_sht_vertical_solve(
        t._plan, L, _vp(d), _vp(th), _vp(ps), _vp(b), _vp(c),
        _vp(m) if m is not None else ctypes.c_void_p(0), _vp(h), _vp(k), float(s_bx), float(s_md),
        float(s_c), float(s_cd), float(s_cn), _vp(d_out), _vp(th_out), _vp(ps_out),
        t.backend.stream()))
    return d_out, th_out, ps_out


def vertical_solve_shifted(transform: SphericalHarmonicTransform, divergence, theta,
                           log_surface_pressure, b_matrix, c_matrix, inverse, k2, tau_w: float):
    """``(I - tau w L) y = r`` for the stage solve of the IMEX integrator
    (reference semi_implicit.py:921-959): forcing = D + tau_w k^2 B x,
    d = inverse_n forcing, x' = x + tau_w C d; ``inverse`` is
    ``helmholtz_inverse(k2, tau_w, 1)``.  Returns projected
    ``(divergence, theta, log_surface_pressure)`` with degree zero passed
    through."""
    return _solve(transform, divergence, theta, log_surface_pressure, b_matrix, c_matrix, None,
                  inverse, k2, tau_w, 0.0, tau_w, 0.0, 1.0)


def vertical_apply(transform: SphericalHarmonicTransform, divergence, theta, log_surface_pressure,
                   b_matrix, c_matrix, structure, inverse, k2, dt_w: float, alpha: float):
    """The off-centred Crank-Nicolson step of the subtracted operator
    (reference semi_implicit.py:887-919): forcing = D + dt k^2 B x -
    alpha (1 - alpha) dt^2 k^2 M D, d = inverse_n forcing, x' = x + dt C
    ((1 - alpha) D + alpha d); ``inverse`` is ``helmholtz_inverse(k2, dt,
    alpha)`` and ``structure`` is M = -B C."""
    a = float(alpha)
    return _solve(transform, divergence, theta, log_surface_pressure, b_matrix, c_matrix,
                  structure, inverse, k2, dt_w, -a * (1.0 - a) * dt_w * dt_w, dt_w, 1.0 - a, a)


def vertical_linear_tendencies(transform: SphericalHarmonicTransform, divergence, theta,
                               log_surface_pressure, b_matrix, c_matrix, k2, weight: float):
    """``w L y``: the divergence, theta and ln ps tendencies of the linear
    operator, each projected (reference semi_implicit.py:796-813)."""
    t = transform
    d, th, ps, L, b, c, (d_out, th_out, ps_out) = _columns(
        t, divergence, theta, log_surface_pressure, b_matrix, c_matrix)
    k = _matrix(t, k2, (t.truncation + 1,), "k2")
    _native.check(t._lib.arwen_sht_vertical_tendencies(
        t._plan, L, _vp(d), _vp(th), _vp(ps), _vp(b), _vp(c), _vp(k), float(weight), _vp(d_out),
        _vp(th_out), _vp(ps_out), t.backend.stream()))
    return d_out, th_out, ps_out


def vertical_half_apply(transform: SphericalHarmonicTransform, divergence, theta,
                        log_surface_pressure, b_matrix, c_matrix, p_d, q_d, p_x, q_x, k2):
    """The square root H = P + Q L of the off-centred Crank-Nicolson map
    (reference semi_implicit.py:845-878): d = P_D D + Q_D (k^2 B x),
    x' = P_x x + Q_x (C D) with the operator's ``symmetric_half_maps``;
    degree zero pinned and projected.  Applied twice it is
    :func:`vertical_apply`."""
    t = transform
    d, th, ps, L, b, c, (d_out, th_out, ps_out) = _columns(
        t, divergence, theta, log_surface_pressure, b_matrix, c_matrix)
    T1 = t.truncation + 1
    pd = _matrix(t, p_d, (T1, L, L), "p_d")
    qd = _matrix(t, q_d, (T1, L, L), "q_d")
    px = _matrix(t, p_x, (T1, L + 1, L + 1), "p_x")
    qx = _matrix(t, q_x, (T1, L + 1, L + 1), "q_x")
    k = _matrix(t, k2, (T1,), "k2")
    _native.check(t._lib.arwen_sht_vertical_half_apply(
        t._plan, L, _vp(d), _vp(th), _vp(ps), _vp(b), _vp(c), _vp(pd), _vp(qd), _vp(px), _vp(qx),
        _vp(k), _vp(d_out), _vp(th_out), _vp(ps_out), t.backend.stream()))
    return d_out, th_out, ps_out


__all__ = ["vertical_apply", "vertical_half_apply", "vertical_linear_tendencies",
           "vertical_solve_shifted", "wave_operator"]
