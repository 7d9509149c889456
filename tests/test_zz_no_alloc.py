"""arwen_sht_sample_static, the sampler entry point that allocates nothing,
against arwen_sht_sample (needs an MI355X).

The static entry point runs the same kernels with the workspace in the
library's own per-device array; its outputs must be bit-identical to the
workspace entry point's, on two calls in a row (the second smaller, so stale
contents of the shared array would show), for values, gradients and winds in
both sampler precisions.  Calls whose workspace exceeds the array are
refused before anything is launched.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle.sht import red_spectrum

pytestmark = pytest.mark.gpu


def _setup(T, count, npts, seed):
    import torch

    from arwen_spectral import SphericalHarmonicTransform

    tr = SphericalHarmonicTransform.create(T, dealias_factor=1.0, backend="hip")
    dev = tr.backend.device()
    c0 = torch.from_numpy(red_spectrum(T, (count,), seed=seed, scale=1.0)
                          .astype(np.complex64)).to(dev)
    c1 = torch.from_numpy(red_spectrum(T, (count,), seed=seed + 1, scale=1.0)
                          .astype(np.complex64)).to(dev)
    rng = np.random.default_rng(seed)
    lat = np.deg2rad(rng.uniform(-89.0, 89.0, npts))
    lon = np.deg2rad(rng.uniform(0.0, 360.0, npts))
    pts = [torch.from_numpy(a).to(dev) for a in (np.sin(lat), np.cos(lat), lon)]
    return tr, c0, c1, pts


def _call(tr, c0, c1, pts, count, npts, kind, precision, static):
    import torch

    from arwen_spectral import _native

    lib = _native.load()
    dev = tr.backend.device()
    vp = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
    outs = [torch.full((count, npts), float("nan"), dtype=torch.float64, device=dev)
            for _ in range(1 if kind == 0 else 2)]
    o1 = outs[1] if kind else None
    second = c1 if kind == 2 else None
    if static:
        rc = lib.arwen_sht_sample_static(tr.plan_handle, vp(c0), vp(second), count, vp(pts[0]),
                                         vp(pts[1]), vp(pts[2]), npts, kind, precision,
                                         vp(outs[0]), vp(o1), tr.backend.stream())
    else:
        nbytes = int(lib.arwen_sht_sample_workspace_bytes(tr.plan_handle, count, kind, npts))
        work = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        rc = lib.arwen_sht_sample(tr.plan_handle, vp(c0), vp(second), count, vp(pts[0]),
                                  vp(pts[1]), vp(pts[2]), npts, kind, precision, vp(work), nbytes,
                                  vp(outs[0]), vp(o1), tr.backend.stream())
    _native.check(rc)
    torch.cuda.synchronize()
    return outs


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("kind", [0, 1, 2])
def test_static_sampler_matches_the_workspace_entry_point(kind, precision):
    import torch

    for T, count, npts in ((63, 12, 3000), (63, 5, 700)):  # the second call is smaller
        tr, c0, c1, pts = _setup(T, count, npts, seed=11 * kind + precision + count)
        want = _call(tr, c0, c1, pts, count, npts, kind, precision, static=False)
        got = _call(tr, c0, c1, pts, count, npts, kind, precision, static=True)
        for g, w in zip(got, want):
            assert torch.equal(g, w), (T, count, npts, kind, precision)
            assert bool(torch.isfinite(g).all())


def test_static_sampler_refuses_what_does_not_fit():
    """A call whose workspace exceeds the static array is refused on the host
    (EINVAL, nothing launched); every buffer passed is sized for the call, so
    only the array's own size stands between it and the refusal."""
    import torch

    from arwen_spectral import _native

    lib = _native.load()
    limit = _native.SAMPLE_STATIC_BYTES
    probe = _setup(63, 8, 64, seed=5)[0]
    fixed = int(lib.arwen_sht_sample_workspace_bytes(probe.plan_handle, 8, 1, 0))
    per_point = int(lib.arwen_sht_sample_workspace_bytes(probe.plan_handle, 8, 1, 1)) - fixed
    too_many = (limit - fixed) // per_point + 1
    assert fixed + too_many * per_point > limit
    tr, c0, c1, pts = _setup(63, 8, too_many, seed=5)
    dev = tr.backend.device()
    outs = [torch.empty((8, too_many), dtype=torch.float64, device=dev) for _ in range(2)]
    vp = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    rc = lib.arwen_sht_sample_static(tr.plan_handle, vp(c0), None, 8, vp(pts[0]), vp(pts[1]),
                                     vp(pts[2]), too_many, 1, 0, vp(outs[0]), vp(outs[1]),
                                     tr.backend.stream())
    assert rc == _native.EINVAL
    assert "static" in lib.arwen_sht_last_error().decode()
    # one point fewer fits and runs
    n = too_many - 1
    got = _call(tr, c0, c1, [p[:n] for p in pts], 8, n, 1, 0, static=True)
    want = _call(tr, c0, c1, [p[:n] for p in pts], 8, n, 1, 0, static=False)
    for g, w in zip(got, want):
        assert torch.equal(g, w)
