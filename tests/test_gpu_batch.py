"""Grouped batches through the C ABI: schedule configs, the autotuner, graph
replays and mixed truncations (needs an MI355X).

A batch is the engine's own API (no reference counterpart); these tests pin
it to the single-operation path and to the FP64 oracle:
  * every column tile and grid gives bit-identical results (each output keeps
    its accumulation order), and so does the autotuner's choice;
  * sharing one accumulator in the split modes stays within the mode's
    tolerance of the oracle, and there the 96-column tile (which has no split
    accumulators) is bit-identical to the 32- and 64-column ones;
  * graph replays, with eager runs between them, match the eager result;
  * a batch mixing T21 and T63 jobs matches the single operations bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle.sht import OracleTransform, red_spectrum

from _helpers import hip_transform, rel_err

pytestmark = pytest.mark.gpu

TOL = {"float32": 1e-6, "bf16x6": 1e-6, "bf16x3": 6e-5}


@pytest.fixture(scope="module")
def oracles():
    return {
        21: OracleTransform.create(21),                       # 33 x 66: Stockham FFT path
        63: OracleTransform.create(63, dealias_factor=1.0),   # 64 x 128: wave-resident FFT path
    }


def _round_trip(oracle, precision, levels=2, seed=0):
    """Device buffers and jobs of a round trip: wind + inverse, then vordiv + forward."""
    import torch
    from arwen_spectral import TransformBatch

    tr = hip_transform(oracle, precision)
    T = oracle.grid.truncation
    dev = torch.device("cuda", torch.cuda.current_device())
    c64 = lambda a: torch.from_numpy(np.ascontiguousarray(a).astype(np.complex64)).to(dev)
    zeta = c64(oracle.project(red_spectrum(T, (levels,), seed=seed + 1, scale=1e-5)))
    div = c64(oracle.project(red_spectrum(T, (levels,), seed=seed + 2, scale=1e-5)))
    scal = c64(oracle.project(red_spectrum(T, (levels + 1,), seed=seed + 3, scale=1.0, mean=0.5)))
    g = lambda n: torch.empty((n, *tr.grid.shape), dtype=torch.float32, device=dev)
    s = lambda n: torch.empty((n, *tr.spectral_shape), dtype=torch.complex64, device=dev)
    bufs = dict(zeta=zeta, div=div, scal=scal, u=g(levels), v=g(levels), grid=g(levels + 1),
                z2=s(levels), d2=s(levels), s2=s(levels + 1))
    jobs = [("wind", tr, (zeta, div), (bufs["u"], bufs["v"])),
            ("inverse", tr, scal, bufs["grid"]),
            ("vordiv", tr, (bufs["u"], bufs["v"]), (bufs["z2"], bufs["d2"])),
            ("forward", tr, bufs["grid"], bufs["s2"])]
    return tr, TransformBatch(jobs), bufs


def _outputs(bufs):
    return {k: bufs[k].clone() for k in ("u", "v", "grid", "z2", "d2", "s2")}


def _run(batch):
    import torch

    batch.run()
    torch.cuda.synchronize()


@pytest.mark.parametrize("precision", ["float32", "bf16x6"])
@pytest.mark.parametrize("T", [21, 63])
def test_schedule_configs_and_autotune_are_bit_identical(oracles, T, precision):
    import torch

    o = oracles[T]
    _, batch, bufs = _round_trip(o, precision)
    assert batch.config() == {"cols": [32, 32], "grid": [0, 0], "split_acc": 1}
    _run(batch)
    ref = _outputs(bufs)
    # the round trip itself against the oracle
    zeta = bufs["zeta"].cpu().numpy()
    ou, ov = o.wind_from_vordiv(zeta, bufs["div"].cpu().numpy())
    assert max(rel_err(ref["u"], ou), rel_err(ref["v"], ov)) < TOL[precision]
    assert rel_err(ref["grid"], o.inverse(bufs["scal"].cpu().numpy())) < TOL[precision]
    wide = ((96, 96), (96, 32)) if precision == "float32" else ()
    for cols in ((64, 64), (64, 32), (32, 64)) + wide:
        for grid in ((0, 0), (64, 1024)):
            batch.set_config(cols=cols, grid=grid)
            _run(batch)
            for k, v in _outputs(bufs).items():
                assert torch.equal(v, ref[k]), (cols, grid, k)
    tuned = batch.autotune(reps=1)
    allowed = (32, 64, 96) if precision == "float32" else (32, 64)
    assert tuned["cols"][0] in allowed and tuned["cols"][1] in allowed
    _run(batch)
    for k, v in _outputs(bufs).items():
        assert torch.equal(v, ref[k]), ("autotuned", tuned, k)


@pytest.mark.parametrize("T", [21, 63])
def test_shared_accumulator_stays_within_tolerance(oracles, T):
    import torch

    o = oracles[T]
    _, batch, bufs = _round_trip(o, "bf16x6")
    batch.set_config(split_acc=0)
    _run(batch)
    zeta, div = bufs["zeta"].cpu().numpy(), bufs["div"].cpu().numpy()
    ou, ov = o.wind_from_vordiv(zeta, div)
    assert max(rel_err(bufs["u"], ou), rel_err(bufs["v"], ov)) < TOL["bf16x6"]
    assert rel_err(bufs["grid"], o.inverse(bufs["scal"].cpu().numpy())) < TOL["bf16x6"]
    oz, od = o.vordiv_from_wind(bufs["u"].cpu().numpy().astype(np.float64),
                                bufs["v"].cpu().numpy().astype(np.float64))
    assert max(rel_err(bufs["z2"], oz), rel_err(bufs["d2"], od)) < TOL["bf16x6"]
    # with one accumulator the 96-column tile joins the others, bit for bit
    ref = _outputs(bufs)
    for cols in ((96, 96), (96, 32), (64, 96)):
        batch.set_config(cols=cols)
        _run(batch)
        for k, v in _outputs(bufs).items():
            assert torch.equal(v, ref[k]), (cols, k)
    tuned = batch.autotune(reps=1)
    assert tuned["split_acc"] == 0 and set(tuned["cols"]) <= {32, 64, 96}
    _run(batch)
    for k, v in _outputs(bufs).items():
        assert torch.equal(v, ref[k]), ("autotuned", tuned, k)


def test_config_rejects_bad_values(oracles):
    _, batch, _ = _round_trip(oracles[21], "float32", levels=1)
    for kw in (dict(cols=(48, 32)), dict(grid=(-1, 0)), dict(split_acc=2)):
        with pytest.raises(ValueError):
            batch.set_config(**kw)
    assert batch.config() == {"cols": [32, 32], "grid": [0, 0], "split_acc": 1}
    # the 96-column tile has no split accumulators: with them on (the default)
    # a split-mode batch refuses it
    _, split, _ = _round_trip(oracles[21], "bf16x6", levels=1)
    with pytest.raises(ValueError, match="split_acc = 0"):
        split.set_config(cols=(96, 32))
    split.set_config(split_acc=0)
    split.set_config(cols=(96, 96))
    assert split.config()["cols"] == [96, 96]


@pytest.mark.parametrize("precision", ["float32", "bf16x6"])
def test_graph_replays_match_eager_runs(oracles, precision):
    import torch

    _, batch, bufs = _round_trip(oracles[63], precision)
    _run(batch)
    ref = _outputs(bufs)
    batch.capture()
    for rep in range(4):
        for k in ref:
            bufs[k].fill_(float("nan") if bufs[k].is_floating_point() else 0)
        torch.cuda.synchronize()
        batch.launch()
        torch.cuda.synchronize()
        for k, v in _outputs(bufs).items():
            assert torch.equal(v, ref[k]), (rep, k)
        _run(batch)  # eager work between replays
        for k, v in _outputs(bufs).items():
            assert torch.equal(v, ref[k]), ("eager", rep, k)


def test_mixed_truncation_batch_matches_single_operations(oracles):
    import torch
    from arwen_spectral import TransformBatch

    dev = torch.device("cuda", torch.cuda.current_device())
    jobs, singles = [], []
    for T in (21, 63):
        o = oracles[T]
        tr = hip_transform(o, "bf16x6")
        coeff = torch.from_numpy(o.project(red_spectrum(T, (3,), seed=T, scale=1.0))
                                 .astype(np.complex64)).to(dev)
        grid = torch.empty((3, *tr.grid.shape), dtype=torch.float32, device=dev)
        back = torch.empty((3, *tr.spectral_shape), dtype=torch.complex64, device=dev)
        jobs += [("inverse", tr, coeff, grid), ("forward", tr, grid, back)]
        singles.append((tr, coeff, grid, back))
    batch = TransformBatch(jobs)
    _run(batch)
    for tr, coeff, grid, back in singles:
        g1 = tr.inverse(coeff)
        b1 = tr.forward(g1)
        torch.cuda.synchronize()
        assert torch.equal(g1, grid)
        assert torch.equal(b1, back)


def test_t383_headline_schedule_matches_the_oracle():
    """The benchmark's schedule at its truncation: one shared accumulator and
    96-column tiles (two levels and three scalars, the T383 768 x 384 grid).
    Within the bf16x6 tolerance of the FP64 oracle, bit-identical to the 32-
    and 64-column tiles, and replayed from a graph bit for bit."""
    import torch

    o = OracleTransform.create(383, dealias_factor=1.0)
    _, batch, bufs = _round_trip(o, "bf16x6", levels=2, seed=7)
    batch.set_config(split_acc=0, cols=(96, 96))
    _run(batch)
    ref = _outputs(bufs)
    zeta, div = bufs["zeta"].cpu().numpy(), bufs["div"].cpu().numpy()
    ou, ov = o.wind_from_vordiv(zeta, div)
    assert max(rel_err(ref["u"], ou), rel_err(ref["v"], ov)) < TOL["bf16x6"]
    assert rel_err(ref["grid"], o.inverse(bufs["scal"].cpu().numpy())) < TOL["bf16x6"]
    og = ref["grid"].cpu().numpy().astype(np.float64)
    assert rel_err(ref["s2"], o.forward(og)) < TOL["bf16x6"]
    oz, od = o.vordiv_from_wind(ref["u"].cpu().numpy().astype(np.float64),
                                ref["v"].cpu().numpy().astype(np.float64))
    ez, ed = rel_err(ref["z2"], oz), rel_err(ref["d2"], od)
    print(f"T383 headline schedule: vorticity {ez:.2e}, divergence {ed:.2e}")
    # the vector analysis weights the winds by m / cos and the derivative
    # table; its 8 lowest orders run in FP64 (fp64_orders), which keeps it
    # within the scalar bound (2.4e-6 / 3.1e-6 without them)
    assert max(ez, ed) < TOL["bf16x6"]
    for cols in ((32, 32), (64, 64), (96, 32)):
        batch.set_config(cols=cols)
        _run(batch)
        for k, v in _outputs(bufs).items():
            assert torch.equal(v, ref[k]), (cols, k)
    batch.set_config(cols=(96, 96))
    batch.capture()
    for rep in range(2):
        for k in ref:
            bufs[k].fill_(float("nan") if bufs[k].is_floating_point() else 0)
        torch.cuda.synchronize()
        batch.launch()
        torch.cuda.synchronize()
        for k, v in _outputs(bufs).items():
            assert torch.equal(v, ref[k]), ("graph", rep, k)


def _member_batch(specs, precision, levels):
    """Round-trip jobs (wind + inverse, then vordiv + forward) for every
    instance of ``specs`` = [(T, instances)], each instance with ``levels``
    levels of vorticity, divergence and temperature plus ln ps; instance k of
    truncation T draws its red spectra from seeds 1000 T + 10 k + {1, 2, 3}."""
    import torch
    from arwen_spectral import SphericalHarmonicTransform

    dev = torch.device("cuda", torch.cuda.current_device())
    c64 = lambda a: torch.from_numpy(np.ascontiguousarray(a).astype(np.complex64)).to(dev)
    parts, jobs = [], []
    for T, count in specs:
        tr = SphericalHarmonicTransform.create(T, dealias_factor=1.0, precision=precision)
        g = lambda n: torch.empty((n, *tr.grid.shape), dtype=torch.float32, device=dev)
        s = lambda n: torch.empty((n, *tr.spectral_shape), dtype=torch.complex64, device=dev)
        for k in range(count):
            seed = 1000 * T + 10 * k
            zeta = red_spectrum(T, (levels,), seed=seed + 1, scale=1e-5)
            div = red_spectrum(T, (levels,), seed=seed + 2, scale=2e-6)
            zeta[..., 0, 0] = 0.0
            div[..., 0, 0] = 0.0
            scal = red_spectrum(T, (levels + 1,), seed=seed + 3, scale=1.0)
            b = dict(T=T, tr=tr, zeta=c64(zeta), div=c64(div), scal=c64(scal),
                     u=g(levels), v=g(levels), grid=g(levels + 1),
                     z2=s(levels), d2=s(levels), s2=s(levels + 1))
            jobs += [("wind", tr, (b["zeta"], b["div"]), (b["u"], b["v"])),
                     ("inverse", tr, b["scal"], b["grid"]),
                     ("vordiv", tr, (b["u"], b["v"]), (b["z2"], b["d2"])),
                     ("forward", tr, b["grid"], b["s2"])]
            parts.append(b)
    return parts, jobs


def _check_against_single_batches(parts, split_acc=0):
    """Every instance of a combined batch is bit-identical to the same jobs
    run as a batch of their own (with the same accumulator split; column
    tiles and grids never change bits)."""
    import torch
    from arwen_spectral import TransformBatch

    for b in parts:
        out = {k: b[k].clone() for k in ("u", "v", "grid", "z2", "d2", "s2")}
        tr = b["tr"]
        alone = TransformBatch([("wind", tr, (b["zeta"], b["div"]), (b["u"], b["v"])),
                                ("inverse", tr, b["scal"], b["grid"]),
                                ("vordiv", tr, (b["u"], b["v"]), (b["z2"], b["d2"])),
                                ("forward", tr, b["grid"], b["s2"])])
        alone.set_config(split_acc=split_acc)
        for k in out:
            b[k].fill_(float("nan"))
        _run(alone)
        for k, v in out.items():
            assert torch.equal(b[k], v), (b["T"], k)


def _check_against_oracle(b, o):
    zeta, div = b["zeta"].cpu().numpy(), b["div"].cpu().numpy()
    scal = b["scal"].cpu().numpy()
    u64, v64 = o.wind_from_vordiv(zeta, div)
    z64, d64 = o.vordiv_from_wind(u64, v64)
    g64 = o.inverse(scal)
    errs = {"u": rel_err(b["u"], u64), "v": rel_err(b["v"], v64), "grid": rel_err(b["grid"], g64),
            "vorticity": rel_err(b["z2"], z64), "divergence": rel_err(b["d2"], d64),
            "scalars": rel_err(b["s2"], o.forward(g64))}
    for k, e in errs.items():
        assert e < TOL["bf16x6"], (b["T"], k, e)
    return errs


def test_ensemble_of_32_members_in_one_launch():
    """C3 at reduced depth: 32 T383 members (4 levels each) in one batch, one
    grouped GEMM launch per direction, autotuned and replayed from a graph.
    Members 0 and 31 match the FP64 oracle within 1e-6; every member is
    bit-identical to its own single-member batch."""
    import torch
    from arwen_spectral import TransformBatch

    parts, jobs = _member_batch([(383, 32)], "bf16x6", levels=4)
    batch = TransformBatch(jobs)
    batch.set_config(split_acc=0)
    batch.autotune(reps=1)
    batch.capture()
    batch.launch()
    torch.cuda.synchronize()
    o = OracleTransform.create(383, dealias_factor=1.0)
    for k in (0, 31):
        errs = _check_against_oracle(parts[k], o)
        print(f"member {k}: " + ", ".join(f"{n} {e:.2e}" for n, e in errs.items()))
    _check_against_single_batches(parts[:2] + parts[-2:])


def test_mixed_truncations_in_one_launch():
    """C4 at reduced depth: 4 T767 + 8 T383 + 16 T127 instances (2 levels
    each) in one batch.  The first and last T127 and T383 instances match the
    FP64 oracle within 1e-6; the T767 instances are bit-identical to a batch
    of their own and their scalar and vector round trips return their inputs
    within 1e-6; every instance checked is bit-identical to its own batch."""
    import torch
    from arwen_spectral import TransformBatch

    parts, jobs = _member_batch([(767, 4), (383, 8), (127, 16)], "bf16x6", levels=2)
    batch = TransformBatch(jobs)
    batch.set_config(split_acc=0)
    batch.autotune(reps=1)
    _run(batch)
    for T in (127, 383):
        o = OracleTransform.create(T, dealias_factor=1.0)
        mine = [b for b in parts if b["T"] == T]
        for b in (mine[0], mine[-1]):
            errs = _check_against_oracle(b, o)
            print(f"T{T}: " + ", ".join(f"{n} {e:.2e}" for n, e in errs.items()))
    for b in (b for b in parts if b["T"] == 767):
        assert rel_err(b["z2"], b["zeta"].cpu().numpy()) < TOL["bf16x6"]
        assert rel_err(b["d2"], b["div"].cpu().numpy()) < TOL["bf16x6"]
        assert rel_err(b["s2"], b["scal"].cpu().numpy()) < TOL["bf16x6"]
    checked = [b for b in parts if b["T"] == 767] + [parts[4], parts[11], parts[12], parts[-1]]
    _check_against_single_batches(checked)
