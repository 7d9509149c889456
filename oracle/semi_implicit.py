"""Float64 restatement of the reference's semi-implicit column solves.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  Follows reference
``semi_implicit.py``: ``VerticalModeSemiImplicit.apply`` (:887-919),
``solve_shifted`` (:921-959) and ``_finish`` (:815-843), and the operator's
``helmholtz_inverse`` (:585-608) for building inverses from a modal
decomposition.  ``synthetic_operator`` makes a well-posed operator for the
tests (there is no hybrid coordinate here): M = V diag(c^2) V^-1 with
positive squared wave speeds, C of full column rank and B = -M C^+, so
M = -B C exactly.
"""
from __future__ import annotations

import numpy as np

from .sht import OracleTransform


def synthetic_operator(nlev, *, seed=0, speeds=(10.0, 330.0)):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((nlev, nlev)))
    v = q @ np.diag(rng.uniform(0.5, 2.0, nlev))
    c2 = np.sort(rng.uniform(speeds[0], speeds[1], nlev))[::-1] ** 2
    m = v @ np.diag(c2) @ np.linalg.inv(v)
    c = rng.standard_normal((nlev + 1, nlev)) / np.sqrt(nlev)
    b = -m @ np.linalg.pinv(c)
    return b, c, m, (v, c2)


def helmholtz_inverse(modal, k2, dt, alpha):
    """semi_implicit.py:585-608: V diag(1 / (1 + (alpha dt)^2 k^2 c^2)) V^-1, identity at k = 0."""
    v, c2 = modal
    scale = (float(alpha) * float(dt)) ** 2 * np.asarray(k2, dtype=np.float64)
    inv = np.einsum("ij,nj,jk->nik", v, 1.0 / (1.0 + scale[:, None] * c2[None, :]),
                    np.linalg.inv(v))
    inv[np.asarray(k2) == 0.0] = np.eye(v.shape[0])
    return inv


def _finish(o: OracleTransform, d_new, x_new, divergence, x):
    d_new[:, 0, 0] = divergence[:, 0, 0]
    x_new[:, 0, 0] = x[:, 0, 0]
    return o.project(d_new), o.project(x_new[:-1]), o.project(x_new[-1])


def solve_shifted(o, divergence, theta, lnps, b, c, inverse, k2, tw):
    """semi_implicit.py:944-959 in float64."""
    d = np.asarray(divergence, dtype=np.complex128)
    x = np.concatenate([np.asarray(theta, dtype=np.complex128),
                        np.asarray(lnps, dtype=np.complex128)[None]], axis=0)
    forcing = d + tw * np.asarray(k2)[None, :, None] * np.einsum("ij,jnm->inm", b, x)
    d_new = np.einsum("nij,jnm->inm", inverse, forcing)
    x_new = x + tw * np.einsum("ij,jnm->inm", c, d_new)
    return _finish(o, d_new, x_new, d, x)


def apply(o, divergence, theta, lnps, b, c, m, inverse, k2, dt, alpha):
    """semi_implicit.py:906-919 in float64."""
    d = np.asarray(divergence, dtype=np.complex128)
    x = np.concatenate([np.asarray(theta, dtype=np.complex128),
                        np.asarray(lnps, dtype=np.complex128)[None]], axis=0)
    k2c = np.asarray(k2)[None, :, None]
    forcing = (d + dt * k2c * np.einsum("ij,jnm->inm", b, x)
               - (alpha * (1.0 - alpha) * dt * dt) * k2c * np.einsum("ij,jnm->inm", m, d))
    d_new = np.einsum("nij,jnm->inm", inverse, forcing)
    x_new = x + dt * np.einsum("ij,jnm->inm", c, (1.0 - alpha) * d + alpha * d_new)
    return _finish(o, d_new, x_new, d, x)


def x_modes(b, c):
    """The (theta, ln ps) block C B's modes (semi_implicit.py:517-537): real
    eigenvalues, values at 1e-9 of the largest set to zero."""
    values, vectors = np.linalg.eig(c @ b)
    values = values.real
    values[np.abs(values) <= 1.0e-9 * np.max(np.abs(values))] = 0.0
    return values, vectors.real


def _half_map_scalars(frequencies, dt, alpha):
    """semi_implicit.py:610-627: sqrt(g) = p + i q omega for the off-centred
    Crank-Nicolson factor g of every mode frequency; omega = 0 gives p = 1,
    q = dt / 2."""
    y = frequencies * float(dt)
    g = (1.0 + 1j * (1.0 - alpha) * y) / (1.0 - 1j * alpha * y)
    root = np.sqrt(g)
    p = root.real
    q = np.where(frequencies > 0.0,
                 root.imag / np.where(frequencies > 0.0, frequencies, 1.0), 0.5 * float(dt))
    return p, q


def symmetric_half_maps(modal, xmodal, k2, dt, alpha):
    """semi_implicit.py:629-671: (P_D, Q_D, P_x, Q_x) per total degree."""
    v, c2 = modal
    xv, xw = xmodal
    k = np.sqrt(np.asarray(k2, dtype=np.float64))
    p_d, q_d = _half_map_scalars(k[:, None] * np.sqrt(c2)[None, :], dt, alpha)
    p_x, q_x = _half_map_scalars(k[:, None] * np.sqrt(-xv)[None, :], dt, alpha)
    vi, wi = np.linalg.inv(v), np.linalg.inv(xw)
    return (np.einsum("ij,nj,jk->nik", v, p_d, vi), np.einsum("ij,nj,jk->nik", v, q_d, vi),
            np.einsum("ij,nj,jk->nik", xw, p_x, wi), np.einsum("ij,nj,jk->nik", xw, q_x, wi))


def half_apply(o, divergence, theta, lnps, b, c, maps, k2):
    """semi_implicit.py:863-878 in float64."""
    p_d, q_d, p_x, q_x = maps
    d = np.asarray(divergence, dtype=np.complex128)
    x = np.concatenate([np.asarray(theta, dtype=np.complex128),
                        np.asarray(lnps, dtype=np.complex128)[None]], axis=0)
    l_d = np.asarray(k2)[None, :, None] * np.einsum("ij,jnm->inm", b, x)
    l_x = np.einsum("ij,jnm->inm", c, d)
    d_new = np.einsum("nij,jnm->inm", p_d, d) + np.einsum("nij,jnm->inm", q_d, l_d)
    x_new = np.einsum("nij,jnm->inm", p_x, x) + np.einsum("nij,jnm->inm", q_x, l_x)
    return _finish(o, d_new, x_new, d, x)


def linear_tendencies(o, divergence, theta, lnps, b, c, k2, w):
    """semi_implicit.py:796-813 in float64: w L y, projected."""
    d = np.asarray(divergence, dtype=np.complex128)
    x = np.concatenate([np.asarray(theta, dtype=np.complex128),
                        np.asarray(lnps, dtype=np.complex128)[None]], axis=0)
    dt_ = w * np.asarray(k2)[None, :, None] * np.einsum("ij,jnm->inm", b, x)
    xt = w * np.einsum("ij,jnm->inm", c, d)
    return o.project(dt_), o.project(xt[:-1]), o.project(xt[-1])
