"""Ensemble member synthesis in one batch (needs an MI355X).

``ensemble_grid_fields`` / ``EnsembleSynthesis`` replace the reference's
member loop in ``grid_fields`` (da/ensemble.py:547-579): every member's u, v,
theta, qv, ln ps and the potentials psi = inverse(inverse_laplacian(zeta)),
chi = the same of D, from one grouped batch.  Checked against the FP64 oracle
per member, bit for bit against the single operations the reference calls
(the inverse Laplacian fused into the pack is the spectral operator's own
arithmetic), and across graph replays.
"""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import pytest

from oracle.sht import OracleTransform, red_spectrum

from _helpers import hip_transform, rel_err

pytestmark = pytest.mark.gpu

TOL = {"float32": 1e-6, "bf16x6": 1e-6}


@pytest.fixture(scope="module")
def oracle():
    return OracleTransform.create(63, dealias_factor=1.0)


def _members(o, R, L, seed=0):
    T = o.grid.truncation
    out = []
    for k in range(R):
        s = seed + 100 * k
        out.append({
            "vorticity": o.project(red_spectrum(T, (L,), seed=s + 1, scale=1e-5)),
            "divergence": o.project(red_spectrum(T, (L,), seed=s + 2, scale=2e-6)),
            "theta": o.project(red_spectrum(T, (L,), seed=s + 3, scale=10.0, mean=300.0)),
            "qv": o.project(red_spectrum(T, (L,), seed=s + 4, scale=1e-3, mean=5e-3)),
            "log_surface_pressure": o.project(red_spectrum(T, (), seed=s + 5, scale=1e-2,
                                                           mean=11.5)),
        })
    return out


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("precision", ["float32", "bf16x6"])
def test_every_member_matches_the_oracle_and_the_single_operations(oracle, precision):
    import torch
    from arwen_spectral import VorticityDivergenceOperator, ensemble_grid_fields

    o = oracle
    tr = hip_transform(o, precision)
    R, L = 4, 3
    host = _members(o, R, L)
    dev = torch.device("cuda", torch.cuda.current_device())
    c64 = lambda a: torch.from_numpy(np.ascontiguousarray(a).astype(np.complex64)).to(dev)
    members = [{k: c64(v) for k, v in m.items()} for m in host]
    got = ensemble_grid_fields(tr, members)
    assert set(got) == {"u", "v", "theta", "qv", "lnps", "psi", "chi"}
    assert tuple(got["u"].shape) == (R, L, *tr.grid.shape)
    assert tuple(got["lnps"].shape) == (R, *tr.grid.shape)
    op = VorticityDivergenceOperator(tr)
    for k in range(R):
        h = {n: _np(members[k][n]).astype(np.complex128) for n in members[k]}
        u, v = o.wind_from_vordiv(h["vorticity"], h["divergence"])
        want = {"u": u, "v": v, "theta": o.inverse(h["theta"]), "qv": o.inverse(h["qv"]),
                "lnps": o.inverse(h["log_surface_pressure"]),
                "psi": o.inverse(o.inverse_laplacian(h["vorticity"])),
                "chi": o.inverse(o.inverse_laplacian(h["divergence"]))}
        for n, w in want.items():
            err = rel_err(got[n][k], w)
            assert err <= TOL[precision], (k, n, err)
        # the reference's own calls, one member at a time
        su, sv = op.wind_from_vordiv(members[k]["vorticity"], members[k]["divergence"])
        single = {"u": su, "v": sv, "theta": tr.inverse(members[k]["theta"]),
                  "lnps": tr.inverse(members[k]["log_surface_pressure"]),
                  "psi": tr.inverse(tr.inverse_laplacian(members[k]["vorticity"])),
                  "chi": tr.inverse(tr.inverse_laplacian(members[k]["divergence"]))}
        for n, t in single.items():
            assert np.array_equal(_np(got[n][k]), _np(t)), (k, n)


def test_state_objects_graph_replays_and_refusals(oracle):
    import torch
    from arwen_spectral import EnsembleSynthesis

    o = oracle
    tr = hip_transform(o, "bf16x6")
    R, L = 3, 2
    dev = torch.device("cuda", torch.cuda.current_device())
    host = _members(o, R, L, seed=7)
    states = [SimpleNamespace(atmosphere=SimpleNamespace(
        **{k: torch.from_numpy(v.astype(np.complex64)).to(dev) for k, v in m.items()}))
        for m in host]
    ens = EnsembleSynthesis(tr, R, L, names=("psi", "theta", "lnps"))
    assert set(ens.inputs) == {"vorticity", "theta", "log_surface_pressure"}
    ens.load(states)
    eager = {n: t.clone() for n, t in ens.run().items()}
    torch.cuda.synchronize()
    for _ in range(3):
        out = ens.run(graph=True)
        torch.cuda.synchronize()
        for n in eager:
            assert np.array_equal(_np(out[n]), _np(eager[n])), n
        ens.run()  # eager work between the replays
        torch.cuda.synchronize()
    for k in range(R):
        err = rel_err(eager["psi"][k], o.inverse(o.inverse_laplacian(
            host[k]["vorticity"].astype(np.complex64).astype(np.complex128))))
        assert err <= 1e-6
    with pytest.raises(ValueError, match="expected 3 members"):
        ens.load(states[:2])
    with pytest.raises(ValueError, match="unknown grid fields"):
        EnsembleSynthesis(tr, R, L, names=("w",))
