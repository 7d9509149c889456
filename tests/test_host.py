"""Host-side checks that need no GPU: the C ABI library, argument validation,
grid sizing and exactness rules (reference spectral/grid.py), backend rules."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "arwen_sht.h"


def _declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(arwen_sht_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from arwen_spectral import _native

    lib = _native.load()
    declared = _declared()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_native.EXPORTS)


def test_plan_create_rejects_bad_descriptors_before_touching_a_device():
    from arwen_spectral import _native

    lib = _native.load()
    plan = ctypes.c_void_p()
    cases = [
        (_native.PlanDesc(0, 10, 20, 6.3e6, _native.FP32, None, None), "truncation"),
        (_native.PlanDesc(10, 5, 22, 6.3e6, _native.FP32, None, None), "nlat=5 is too small"),
        (_native.PlanDesc(10, 11, 21, 6.3e6, _native.FP32, None, None), "nlon must be even"),
        (_native.PlanDesc(10, 11, 22, -1.0, _native.FP32, None, None), "radius_m"),
        (_native.PlanDesc(10, 11, 22, 6.3e6, 7, None, None), "precision"),
    ]
    for desc, msg in cases:
        rc = lib.arwen_sht_plan_create(ctypes.byref(desc), ctypes.byref(plan))
        assert rc == _native.EINVAL
        assert msg in lib.arwen_sht_last_error().decode()
        with pytest.raises(ValueError, match=re.escape(msg)):
            _native.check(rc)
    assert lib.arwen_sht_batch_create(None, 0, ctypes.byref(plan)) == _native.EINVAL


def _grid(truncation, nlat, nlon):
    from arwen_spectral import GaussianGrid

    mu, w = np.polynomial.legendre.leggauss(nlat)
    return GaussianGrid(truncation, nlat, nlon, mu, np.arcsin(mu),
                        np.arange(nlon) * 2 * np.pi / nlon, w, np.sqrt(1 - mu * mu))


def test_grid_shape_rules_match_the_oracle():
    from arwen_spectral import GaussianGrid
    from oracle.sht import grid_shape

    for t in (1, 5, 21, 63, 127, 383, 767):
        for f in (1.0, 1.5, 2.0):
            assert GaussianGrid.shape_for(t, dealias_factor=f) == grid_shape(t, dealias_factor=f)
    assert GaussianGrid.shape_for(383, dealias_factor=1.0) == (384, 768)
    with pytest.raises(ValueError, match="explicit nlon must be even"):
        GaussianGrid.shape_for(10, nlon=23)
    with pytest.raises(ValueError, match="too small"):
        GaussianGrid.shape_for(10, nlat=5)
    with pytest.raises(ValueError, match="truncation must be >= 1"):
        GaussianGrid.shape_for(0)
    with pytest.raises(ValueError, match="truncation must be an integer"):
        GaussianGrid.shape_for(True)


def test_native_grid_exactness_and_aliasing_refusal():
    # reference tests/test_global_spectral_transform.py:222-262
    g = _grid(21, 22, 64)
    assert g.exact_analysis_truncation == 12
    with pytest.raises(ValueError, match=r"31 \+ 21 = 52 > 43") as caught:
        g.truncated_to(21)
    assert "T_in + T <= 2*nlat - 1" in str(caught.value)
    assert "T <= 12 is exact" in str(caught.value)
    assert g.truncated_to(12).truncation == 12
    with pytest.raises(ValueError, match=r"m <= nlon/2 - 1 = 31"):
        g.truncated_to(32)
    native = _grid(1535, 1536, 3072)
    assert native.zonal_content_truncation == 1535
    assert native.exact_analysis_truncation == 1535


def test_backend_rules():
    from arwen_spectral import get_backend

    with pytest.raises(ValueError, match="backend must be 'hip'"):
        get_backend("numpy")
    with pytest.raises(ValueError, match="precision must be one of"):
        get_backend("hip", "float16")


def test_global_mean_uses_the_quadrature():
    g = _grid(15, 16, 32)
    field = np.broadcast_to((3 * g.sin_lat ** 2)[:, None], g.shape)
    assert abs(g.global_mean(field) - 1.0) < 1e-14
    assert abs(g.global_integral(np.ones(g.shape)) - 4 * np.pi) < 1e-13
