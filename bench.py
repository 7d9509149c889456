#!/usr/bin/env python3
"""Spherical-harmonic round trips per second on one MI355X.

Default workload (the headline metric): T383 on the 768 x 384 linear Gaussian
grid, 121 fields -- 40 levels of vorticity and divergence, 40 of temperature
and ln ps.  One step is one round trip:

    synthesis  vorticity/divergence -> u, v ; temperature, ln ps -> grid
    analysis   u, v -> vorticity/divergence ; temperature, ln ps -> spectral

run as one TransformBatch (one grouped Legendre GEMM launch per direction)
replayed from a hipGraph, inputs resident in HBM.  ``value`` is the split-BF16
(bf16x6) round trips/s; the same workload in FP32 MFMA is timed beside it.

    python bench.py [--steps K] [--warmup W] [--workload t383|t383x32|mixed|sample|vertical]
                    [--precision bf16x6|bf16x3|float32] [--no-cpu-baseline]

Prints one JSON line.  Only ``cpu_baseline()`` touches the FP64 oracle: it
times the oracle (the reference's numpy algorithm) on a bounded sample of the
same fields on this host's cores and returns its outputs for the error check.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "gpuwm-global_amd"))

import numpy as np  # noqa: E402

METRIC = "T383 121-field SH round trips/s: split-BF16 MFMA vs FP32, rel err ≤1e-6"
MFMA_PEAK = {"bf16x6": 2516.6, "bf16x3": 2516.6, "float32": 157.3}   # dense TFLOP/s
PRODUCTS = {"bf16x6": 6, "bf16x3": 3, "float32": 1}
HBM_PEAK_GBS = 8000.0
LEVELS = 40
CLASS_SCALE = {"vorticity": 1e-5, "divergence": 2e-6, "temperature": 10.0, "lnps": 1e-2}


def red_spectrum(truncation, lead, seed, scale):
    """Seeded red-spectrum coefficients: scale * n^-1.5 * (N + iN)/sqrt(2),
    1 <= n <= T, m <= n, m = 0 real, zero above the triangle and at n = 0
    (temperature and ln ps are anomalies about fixed reference values)."""
    t1 = truncation + 1
    rng = np.random.Generator(np.random.PCG64(seed))
    re = rng.standard_normal((*lead, t1, t1), dtype=np.float64)
    im = rng.standard_normal((*lead, t1, t1), dtype=np.float64)
    n = np.arange(t1, dtype=np.float64)
    amp = np.zeros(t1)
    amp[1:] = n[1:] ** -1.5
    c = (re + 1j * im) * (scale * amp / math.sqrt(2.0))[:, None]
    c *= np.tri(t1, dtype=bool)
    c[..., :, 0] = c[..., :, 0].real
    return c


def member_fields(truncation, members, seed0=0):
    """(vorticity, divergence, scalars) per member; member 0 is the control,
    member k adds a 1% seeded red-spectrum perturbation."""
    out = {}
    seeds = {"vorticity": 11, "divergence": 12, "temperature": 13, "lnps": 14}
    for name, s in seeds.items():
        lev = 1 if name == "lnps" else LEVELS
        ctrl = red_spectrum(truncation, (lev,), seed0 + s, CLASS_SCALE[name])
        mem = [ctrl]
        for k in range(1, members):
            mem.append(ctrl + 0.01 * red_spectrum(truncation, (lev,), seed0 + 1000 * k + s,
                                                  CLASS_SCALE[name]))
        out[name] = np.stack(mem)  # (members, lev, T+1, T+1)
    return out


class RoundTrip:
    """The device buffers and the batch of one round-trip workload."""

    def __init__(self, specs, precision, torch, tune=True, split_acc=None, fp64_orders=None):
        from arwen_spectral import SphericalHarmonicTransform, TransformBatch

        self.parts = []
        jobs = []
        dev = torch.device("cuda", torch.cuda.current_device())
        for spec in specs:
            tr = SphericalHarmonicTransform.create(spec["T"], dealias_factor=1.0, precision=precision,
                                                   fp64_orders=fp64_orders)
            f = spec["fields"]
            c64 = lambda a: torch.from_numpy(a.astype(np.complex64)).to(dev)
            zeta = c64(f["vorticity"]).contiguous()
            div = c64(f["divergence"]).contiguous()
            scal = c64(np.concatenate([f["temperature"], f["lnps"]], axis=1)).contiguous()
            M = zeta.shape[0]
            g = lambda n: torch.empty((M, n, *tr.grid.shape), dtype=torch.float32, device=dev)
            s = lambda n: torch.empty((M, n, *tr.spectral_shape), dtype=torch.complex64, device=dev)
            u, v, grid = g(LEVELS), g(LEVELS), g(LEVELS + 1)
            z2, d2, s2 = s(LEVELS), s(LEVELS), s(LEVELS + 1)
            jobs += [("wind", tr, (zeta, div), (u, v)), ("inverse", tr, scal, grid),
                     ("vordiv", tr, (u, v), (z2, d2)), ("forward", tr, grid, s2)]
            self.parts.append(dict(tr=tr, zeta=zeta, div=div, scal=scal, u=u, v=v, grid=grid,
                                   z2=z2, d2=d2, s2=s2, spec=spec))
        self.batch = TransformBatch(jobs)
        if split_acc is not None:
            self.batch.set_config(split_acc=split_acc)
        # the autotuner picks each direction's column tile and grid by timing them
        self.config = self.batch.autotune(reps=3) if tune else self.batch.config()
        self.flops, self.fft_bytes = self.batch.work()


def time_steps(rt, steps, warmup, torch):
    rt.batch.capture()
    for _ in range(warmup):
        rt.batch.launch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        rt.batch.launch()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def class_errors(rt):
    """Round-trip error per field class, max|c' - c| / max|c|, over every field
    of the first and last member of each truncation."""
    out = {}
    for part in rt.parts:
        M = part["zeta"].shape[0]
        sel = sorted({0, M - 1})
        pairs = {"vorticity": (part["z2"], part["zeta"]), "divergence": (part["d2"], part["div"]),
                 "temperature": (part["s2"][:, :LEVELS], part["scal"][:, :LEVELS]),
                 "lnps": (part["s2"][:, LEVELS:], part["scal"][:, LEVELS:])}
        for name, (have, want) in pairs.items():
            h = have[sel].cpu().numpy().astype(np.complex128)
            w = part["spec"]["fields"][name][sel]
            e = float(np.max(np.abs(h - w)) / np.max(np.abs(w)))
            key = f"T{part['spec']['T']}_{name}"
            out[key] = max(out.get(key, 0.0), e)
    return out


def cpu_baseline(part, sample_levels=LEVELS):
    """FP64 oracle (reference numpy algorithm) on the same workload: by default
    all 121 fields of one member (about 5 s per round trip on 16 cores), or the
    first `sample_levels` levels of each class plus ln ps, scaled linearly.
    Returns the timing entry and the oracle's outputs."""
    sys.path.insert(0, str(ROOT))
    from oracle.sht import OracleTransform

    try:
        from threadpoolctl import threadpool_limits
    except ImportError:  # pragma: no cover
        threadpool_limits = None
    cores = min(len(os.sched_getaffinity(0)), 16)
    T = part["spec"]["T"]
    f = part["spec"]["fields"]
    L = sample_levels
    zeta, div = f["vorticity"][0, :L], f["divergence"][0, :L]
    scal = np.concatenate([f["temperature"][0, :L], f["lnps"][0]], axis=0)
    ctx = threadpool_limits(cores) if threadpool_limits else None
    try:
        o = OracleTransform.create(T, dealias_factor=1.0)
        times = []
        for rep in range(3):  # first call, then two steady calls
            t0 = time.perf_counter()
            u, v = o.wind_from_vordiv(zeta, div)
            g = o.inverse(scal)
            z2, d2 = o.vordiv_from_wind(u, v)
            s2 = o.forward(g)
            times.append(time.perf_counter() - t0)
    finally:
        if ctx is not None:
            ctx.__exit__(None, None, None)
    fields_sample = 3 * L + 1
    steady = float(np.mean(times[1:]))
    fields_total = 3 * LEVELS + 1
    value = (fields_sample / fields_total) / steady  # 121-field round trips/s
    what = (f"all {fields_total} fields of one member" if fields_sample == fields_total else
             f"{fields_sample} of {fields_total} fields ({L} levels of vorticity, divergence, "
             f"temperature + ln ps), scaled linearly to {fields_total} fields")
    entry = {"value": value, "unit": "round trips/s", "cores": cores, "kind": "port",
             "sample": f"FP64 numpy oracle (reference per-order BLAS algorithm), T{T}, {what}; "
                       f"steady mean of 2 round trips ({steady:.2f} s each) after a first call "
                       f"({times[0]:.2f} s)"}
    return entry, dict(u=u, v=v, grid=g, zeta=z2, div=d2, scal=s2, L=L)


def oracle_errors(part, ref):
    L = ref["L"]
    def e(have, want):
        h = have.cpu().numpy().astype(np.float64 if np.isrealobj(want) else np.complex128)
        return float(np.max(np.abs(h - want)) / np.max(np.abs(want)))
    return {
        "u": e(part["u"][0, :L], ref["u"]), "v": e(part["v"][0, :L], ref["v"]),
        "temperature_grid": e(part["grid"][0, :L], ref["grid"][:L]),
        "lnps_grid": e(part["grid"][0, LEVELS:], ref["grid"][L:]),
        "vorticity": e(part["z2"][0, :L], ref["zeta"]), "divergence": e(part["d2"][0, :L], ref["div"]),
        "temperature": e(part["s2"][0, :L], ref["scal"][:L]),
        "lnps": e(part["s2"][0, LEVELS:], ref["scal"][L:]),
    }


MODE_ID = {"float32": 1, "bf16x3": 2, "bf16x6": 3}
TRAFFIC_FILE = ROOT / "profiles" / "legendre_gemm_t383_traffic.json"


def gemm_traffic(config, precision, workload):
    """HBM bytes of the round trip's two GEMM launches (synthesis + analysis)
    from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary, each counter
    scaled by its copy-kernel calibration factor; None when the summary does
    not cover this workload and kernel instantiation."""
    if workload != "t383" or not TRAFFIC_FILE.exists():
        return None
    rec = json.loads(TRAFFIC_FILE.read_text())
    total = 0.0
    for d in (0, 1):
        acc2 = "true" if (precision != "float32" and config["split_acc"]) else "false"
        name = f"legendre_gemm_kernel<{MODE_ID[precision]}, {d}, {config['cols'][d]}, {acc2}>"
        k = rec["kernels"].get(name)
        if k is None:
            return None
        total += (k["fetch_kib"] * rec["fetch_factor"] + k["write_kib"] * rec["write_factor"]) * 1024
    return total


TABLE_BYTES = {"float32": 4, "bf16x3": 4, "bf16x6": 6}  # bytes per table element the GEMM reads


def gemm_algorithmic_bytes(rt, precision):
    """Unique HBM bytes of the round trip's two GEMM launches: each Legendre
    table the launch uses read once (P, dP/dphi, P m/cos: K_tri x nlat/2
    folded elements, no padding), and every field's coefficients (K_tri
    complex64) and waist (nlat x (T+1) complex64) read or written once per
    direction."""
    total = 0.0
    for part in rt.parts:
        T, nlat = part["spec"]["T"], part["tr"].grid.nlat
        members = part["zeta"].shape[0]
        k_tri = (T + 1) * (T + 2) // 2
        tables = 3 * k_tri * (nlat / 2) * TABLE_BYTES[precision]
        fields = members * (3 * LEVELS + 1)
        data = fields * (8.0 * k_tri + 8.0 * nlat * (T + 1))
        total += 2 * (tables + data)
    return total


def roofline(rt, phases, precision, workload="t383"):
    """Roofline of the dominant kernel.  For the Legendre GEMM the bound is
    chosen by arithmetic intensity against the ridge point: MFMA products per
    algorithmic byte above peak_mfma / peak_hbm is MFMA-bound, below it
    HBM-bound; the other roof is reported beside it."""
    gemm_ms = phases["synthesis_gemm"] + phases["analysis_gemm"]
    fft_ms = phases["inverse_fft"] + phases["forward_fft"]
    s = PRODUCTS[precision]
    if gemm_ms >= fft_ms:
        t = gemm_ms * 1e-3
        alg_bytes = gemm_algorithmic_bytes(rt, precision)
        intensity = s * rt.flops / alg_bytes
        ridge = MFMA_PEAK[precision] * 1e12 / (HBM_PEAK_GBS * 1e9)
        mfma = {"bound": "mfma", "achieved": s * rt.flops / t / 1e12,
                "peak": MFMA_PEAK[precision], "unit": "TFLOP/s"}
        hbm = {"bound": "hbm", "achieved": alg_bytes / t / 1e9, "peak": HBM_PEAK_GBS,
               "unit": "GB/s"}
        for r in (mfma, hbm):
            r["frac"] = r["achieved"] / r["peak"]
        main, other = (mfma, hbm) if intensity >= ridge else (hbm, mfma)
        return {**main, "kernel": "legendre_gemm (synthesis + analysis launches)",
                "traffic": gemm_traffic(rt.config, precision, workload),
                "traffic_unit": "HBM bytes per round trip (both launches)",
                "algorithmic_bytes": alg_bytes, "intensity_flop_per_byte": intensity,
                "ridge_flop_per_byte": ridge, "other_roof": other,
                "fp32_equivalent_tflops": rt.flops / t / 1e12, "mfma_products_per_fma": s}
    achieved = rt.fft_bytes / (fft_ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": "rfft_fold + irfft_unfold", "achieved": achieved,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None}


SAMPLE_T, SAMPLE_FIELDS, SAMPLE_POINTS = 383, 40, 30000
FP64_PEAK_TFLOPS = 78.6


def sampler_bench(args, torch):
    """Point sampler (sample_scalar): T383, 40 fields (one level stack) at
    30,000 scattered points, inputs resident on the device; timed with HIP
    events on the launch stream, float64 (`value`) and float32 beside it.
    Algorithmic work per call: per point and packed (n, m), 4 FLOP per field
    (a complex coefficient times a real basis value) plus 4 FLOP of FP64
    recurrence; the bound is the FP64 vector peak."""
    import ctypes

    from arwen_spectral import SphericalHarmonicTransform, _native

    T, F, N = SAMPLE_T, SAMPLE_FIELDS, SAMPLE_POINTS
    t = SphericalHarmonicTransform.create(T, dealias_factor=1.0, backend="hip")
    dev = torch.device("cuda", torch.cuda.current_device())
    coeff = red_spectrum(T, (F,), 31, CLASS_SCALE["temperature"])
    c = torch.from_numpy(coeff.astype(np.complex64)).to(dev)
    rng = np.random.Generator(np.random.PCG64(32))
    lat = np.deg2rad(rng.uniform(-90.0, 90.0, N))
    lon = np.deg2rad(rng.uniform(0.0, 360.0, N))
    x = torch.from_numpy(np.sin(lat)).to(dev)
    cl = torch.from_numpy(np.cos(lat)).to(dev)
    lo = torch.from_numpy(lon).to(dev)
    lib = _native.load()
    nbytes = int(lib.arwen_sht_sample_workspace_bytes(t.plan_handle, F, 0, N))
    work = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    out = torch.empty((F, N), dtype=torch.float64, device=dev)
    vp = lambda a: ctypes.c_void_p(a.data_ptr())  # noqa: E731

    def call(prec):
        _native.check(lib.arwen_sht_sample(t.plan_handle, vp(c), None, F, vp(x), vp(cl), vp(lo), N,
                                           0, prec, vp(work), nbytes, vp(out), None,
                                           t.backend.stream()))

    k_tri = (T + 1) * (T + 2) // 2
    flops = float(N) * k_tri * (4.0 * F + 4.0)
    res = {}
    for prec, name in ((0, "float64"), (1, "float32")):
        for _ in range(args.warmup):
            call(prec)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(args.steps):
            call(prec)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        res[name] = {"ms_per_call": ms, "value": F * N / (ms * 1e-3),
                     "sample_max": float(out.abs().max().item())}
        print(f"[bench] sampler {name}: {ms:.3f} ms per call", file=sys.stderr, flush=True)
    # the other kinds at float64: gradients of the same 40 fields, winds of
    # 20 vorticity/divergence levels (40 psi/chi columns)
    out1 = torch.empty((F, N), dtype=torch.float64, device=dev)
    nb = int(lib.arwen_sht_sample_workspace_bytes(t.plan_handle, F, 2, N))
    if nb > nbytes:
        work = torch.empty(nb, dtype=torch.uint8, device=dev)
        nbytes = nb
    for kind, name, count, second in ((1, "gradient", F, None), (2, "wind", F // 2, c[F // 2:])):
        def call_kind():
            _native.check(lib.arwen_sht_sample(
                t.plan_handle, vp(c), vp(second) if second is not None else None, count, vp(x),
                vp(cl), vp(lo), N, kind, 0, vp(work), nbytes, vp(out), vp(out1),
                t.backend.stream()))
        call_kind()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(args.steps):
            call_kind()
        e1.record()
        torch.cuda.synchronize()
        kms = e0.elapsed_time(e1) / args.steps
        res[name] = {"ms_per_call": kms, "value": count * N / (kms * 1e-3), "outputs": count}
        print(f"[bench] sampler {name} (float64): {kms:.3f} ms per call", file=sys.stderr,
              flush=True)
    ms = res["float64"]["ms_per_call"]
    achieved = flops / (ms * 1e-3) / 1e12
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, str(ROOT))
        from oracle import sampling as osamp
        from oracle.sht import OracleTransform

        o = OracleTransform.create(T, dealias_factor=1.0)
        sub = 200
        lat_d, lon_d = np.rad2deg(lat[:sub]), np.rad2deg(lon[:sub])
        t0 = time.perf_counter()
        osamp.sample_scalar(o, coeff.astype(np.complex64).astype(np.complex128), lat_d, lon_d)
        sec = time.perf_counter() - t0
        cpu = {"value": F * sub / sec, "unit": "point samples/s", "cores": 1, "kind": "port",
               "sample": f"FP64 numpy oracle of the reference's host sampler "
                         f"(sampling.py:382-426), T{T}, {F} fields at {sub} of the {N} points"}
    print(json.dumps({
        "metric": f"T{T} point samples/s ({F} fields at {N} scattered points, sample_scalar)",
        "value": res["float64"]["value"], "unit": "point samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded red-spectrum coefficients, uniform random points",
        "config": {"workload": f"sample_scalar T{T}, {F} fields, {N} points", "T": T,
                   "fields": F, "points": N},
        "float32": res["float32"], "gradient": res["gradient"], "wind": res["wind"],
        "roofline": {"bound": "valu_fp64",
                     "kernel": "sample_pack + sample_kernel<8, 0, false> + sample_reduce",
                     "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "algorithmic_flop_per_call": flops},
        "cpu_baseline": cpu}))


def vertical_bench(args, torch):
    """Semi-implicit stage solve on spectral columns (solve_shifted): T383,
    40 levels of divergence and theta plus ln ps, synthetic well-posed
    vertical operator, inputs resident on the device; HIP events on the
    launch stream.  Algorithmic bytes per call: the triangles of the 2L + 1
    input planes read once (complex64), the 2L + 1 output planes written
    once (whole (T+1)^2 squares: the outputs are fresh arrays, zero above
    the triangle), the per-degree inverses read once."""
    import ctypes  # noqa: F401

    from arwen_spectral import SphericalHarmonicTransform
    from arwen_spectral.semi_implicit import vertical_solve_shifted, wave_operator

    sys.path.insert(0, str(ROOT))
    from oracle import semi_implicit as osi

    T, L = 383, LEVELS
    t = SphericalHarmonicTransform.create(T, dealias_factor=1.0, backend="hip")
    dev = torch.device("cuda", torch.cuda.current_device())
    b, c, m, modal = osi.synthetic_operator(L, seed=7)
    k2 = wave_operator(t)
    tw = 900.0
    inv = osi.helmholtz_inverse(modal, k2, tw, 1.0)
    f32 = lambda a: torch.from_numpy(np.asarray(a, dtype=np.float32)).to(dev)  # noqa: E731
    c64 = lambda a: torch.from_numpy(a.astype(np.complex64)).to(dev)  # noqa: E731
    d = c64(red_spectrum(T, (L,), 41, CLASS_SCALE["divergence"]))
    th = c64(red_spectrum(T, (L,), 42, CLASS_SCALE["temperature"]))
    ps = c64(red_spectrum(T, (), 43, CLASS_SCALE["lnps"]))
    bm, cm, im, km = f32(b), f32(c), f32(inv), f32(k2)
    for _ in range(args.warmup):
        vertical_solve_shifted(t, d, th, ps, bm, cm, im, km, tw)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(args.steps):
        vertical_solve_shifted(t, d, th, ps, bm, cm, im, km, tw)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    k_tri = (T + 1) * (T + 2) // 2
    planes = 2 * L + 1
    alg = planes * (8.0 * k_tri + 8.0 * (T + 1) ** 2) + 4.0 * (T + 1) * L * L
    achieved = alg / (ms * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        from oracle.sht import OracleTransform

        o = OracleTransform.create(T, dealias_factor=1.0)
        dh, thh, psh = (x.cpu().numpy() for x in (d, th, ps))
        t0 = time.perf_counter()
        osi.solve_shifted(o, dh, thh, psh, b, c, inv, k2, tw)
        sec = time.perf_counter() - t0
        cpu = {"value": 1.0 / sec, "unit": "solves/s", "cores": min(len(os.sched_getaffinity(0)), 16),
               "kind": "port", "sample": f"FP64 numpy oracle of solve_shifted (semi_implicit.py:"
                                        f"944-959), T{T}, {L} levels, one call"}
    print(json.dumps({
        "metric": f"T{T} L{L} semi-implicit stage solves/s (solve_shifted on spectral columns)",
        "value": 1e3 / ms, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: seeded red-spectrum fields, synthetic well-posed vertical operator",
        "config": {"workload": f"solve_shifted T{T}, {L} levels", "T": T, "levels": L},
        "roofline": {"bound": "hbm", "kernel": "vertical_solve_kernel", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "algorithmic_bytes": alg},
        "cpu_baseline": cpu}))


def workload_specs(name):
    if name == "t383":
        return [dict(T=383, members=1)]
    if name == "t383x32":
        return [dict(T=383, members=32)]
    if name == "mixed":
        return [dict(T=767, members=4), dict(T=383, members=8), dict(T=127, members=16)]
    raise SystemExit(f"unknown workload {name}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="t383")
    ap.add_argument("--precision", default="bf16x6")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--split-acc", type=int, default=0,
                    help="split modes: 1 = small products in their own accumulators; 0 (the "
                         "default) = one accumulator, which admits the 96-column GEMM tile and "
                         "keeps the FP32 path's error class")
    ap.add_argument("--fp64-orders", type=int, default=None,
                    help="vector operators: zonal orders below this run in FP64 end to end "
                         "(default: 0 for float32, 4 ceil((T+1)/96) for the split modes)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB on the device three times (the known-size kernel "
                         "that scales rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the transform "
              "would shard by members or fields, with no exchange between GPUs)", file=sys.stderr)
        sys.exit(2)
    import torch

    torch.cuda.set_device(0)
    if args.copy_calibration:
        n = 64 * 1024 * 1024
        x = torch.ones(n, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 4, "copies": 3}))
        return
    if args.workload == "sample":
        sampler_bench(args, torch)
        return
    if args.workload == "vertical":
        vertical_bench(args, torch)
        return
    specs = workload_specs(args.workload)
    for sp in specs:
        sp["fields"] = member_fields(sp["T"], sp["members"])
    nfields = sum(sp["members"] * (3 * LEVELS + 1) for sp in specs)
    print(f"[bench] workload {args.workload}: {nfields} fields, precision {args.precision}",
          file=sys.stderr, flush=True)

    rt = RoundTrip(specs, args.precision, torch, not args.no_autotune, args.split_acc,
                   args.fp64_orders)
    sec = time_steps(rt, args.steps, args.warmup, torch)
    phases = rt.batch.profile(reps=3)
    errors = class_errors(rt)
    print(f"[bench] {args.precision}: {sec * 1e3:.3f} ms/step, phases {phases}, "
          f"gemm config {rt.config}", file=sys.stderr, flush=True)

    fp32 = None
    if not args.no_fp32 and args.precision != "float32":
        rt32 = RoundTrip(specs, "float32", torch, not args.no_autotune)
        sec32 = time_steps(rt32, args.steps, args.warmup, torch)
        ph32 = rt32.batch.profile(reps=3)
        fp32 = {"value": 1.0 / sec32, "ms_per_step": sec32 * 1e3, "phases_ms": ph32,
                "gemm_config": rt32.config,
                "roundtrip_error": class_errors(rt32),
                "roofline": roofline(rt32, ph32, "float32", args.workload)}
        print(f"[bench] float32: {sec32 * 1e3:.3f} ms/step", file=sys.stderr, flush=True)
        del rt32

    cpu = None
    oracle_err = None
    if not args.no_cpu_baseline:
        part = next(p for p in rt.parts if p["spec"]["T"] == 383) if any(
            p["spec"]["T"] == 383 for p in rt.parts) else rt.parts[0]
        cpu, ref = cpu_baseline(part)
        oracle_err = oracle_errors(part, ref)
        print(f"[bench] cpu baseline {cpu['value']:.4f} round trips/s on {cpu['cores']} cores",
              file=sys.stderr, flush=True)

    members = sum(sp["members"] for sp in specs)
    result = {
        "metric": METRIC,
        "value": 1.0 / sec,
        "unit": "round trips/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"bf16x6": "bf16x6 (three bf16 pieces, fp32 accumulate)",
                  "bf16x3": "bf16x3 (bf16 hi+lo, fp32 accumulate)",
                  "float32": "fp32"}[args.precision],
        "data": "synthetic: seeded red-spectrum coefficients (n^-1.5), temperature and ln ps as "
                "anomalies about fixed references",
        "config": {"workload": "T383 768x384 linear Gaussian grid, 121 fields per member (40 "
                               "vorticity, 40 divergence, 40 temperature, ln ps): synthesis "
                               "incl. u,v from vorticity/divergence, then analysis incl. "
                               "vorticity/divergence from u,v" if args.workload == "t383"
                   else args.workload,
                   "members": members, "fields": nfields, "precision": args.precision,
                   "specs": [{k: v for k, v in sp.items() if k != "fields"} for sp in specs]},
        "phases_ms": phases,
        "gemm_config": rt.config,
        "roundtrip_error": errors,
        "error_vs_fp64_oracle": oracle_err,
        "fp32": fp32,
        "speedup_vs_fp32": (fp32["ms_per_step"] / (sec * 1e3)) if fp32 else None,
        "legendre_gflop_per_step": rt.flops / 1e9,
        "fft_mb_per_step": rt.fft_bytes / 1e6,
        "roofline": roofline(rt, phases, args.precision, args.workload),
        "cpu_baseline": cpu,
    }
    print(json.dumps(result))


if __name__ == "__main__":
    main()
