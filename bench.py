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

    python bench.py [--steps K] [--warmup W] [--precision bf16x6|bf16x3|float32]
                    [--workload t383|t383x32|mixed|t63|step|stepx32|sample|vertical|ensemble]
                    [--no-cpu-baseline] [--no-extras]

``step`` is the dycore's per-step set at T383 (inverse: u,v, temperature with
both gradients, q = exp(ln q), ln ps with both gradients; forward: vorticity
and divergence from u,v, temperature, q, geopotential, ln ps) in one batch;
``stepx32`` is 32 members of it.  The default run also times single calls at
the dycore's chunk shapes (6 x 40 fields, one ln ps plane) and the T63
round trip beside the oracle's T63 figure.

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
FP64_PEAK_TFLOPS = 78.6  # dense FP64 (vector and MFMA)
HBM_PEAK_GBS = 8000.0
LEVELS = 40
CLASS_SCALE = {"vorticity": 1e-5, "divergence": 2e-6, "temperature": 10.0, "lnps": 1e-2,
               "lnq": 0.3, "geopotential": 3e3}


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


def member_fields(truncation, members, seed0=0, step=False):
    """Spectral fields per member: vorticity, divergence, temperature, ln ps
    (and, for the per-step set, ln q and geopotential); member 0 is the
    control, member k adds a 1% seeded red-spectrum perturbation."""
    out = {}
    seeds = {"vorticity": 11, "divergence": 12, "temperature": 13, "lnps": 14}
    if step:
        seeds.update(lnq=15, geopotential=16)
    for name, s in seeds.items():
        lev = 1 if name == "lnps" else LEVELS
        ctrl = red_spectrum(truncation, (lev,), seed0 + s, CLASS_SCALE[name])
        mem = [ctrl]
        for k in range(1, members):
            mem.append(ctrl + 0.01 * red_spectrum(truncation, (lev,), seed0 + 1000 * k + s,
                                                  CLASS_SCALE[name]))
        out[name] = np.stack(mem)  # (members, lev, T+1, T+1)
    return out


#: coefficient planes read and waist planes written by the synthesis GEMM
#: (per level), and waist planes read and coefficient planes written by the
#: analysis GEMM, per job kind
GEMM_PLANES = {"wind": (2, 2, 0, 0), "inverse": (1, 1, 0, 0), "inverse_gradient": (1, 2, 0, 0),
               "vordiv": (0, 0, 2, 2), "forward": (0, 0, 1, 1)}
# Legendre tables each job's GEMM segments read (direction, kinds): P the
# basis, D = dP/dphi, M = P m / cos.  The wind and vordiv jobs of a plan with
# FP64 low orders run their other orders in the recurrence form (the basis
# alone, one degree past the truncation).
GEMM_TABLES = {"wind": ("syn", "DM"), "inverse": ("syn", "P"), "inverse_gradient": ("syn", "PD"),
               "vordiv": ("ana", "MD"), "forward": ("ana", "P")}


def recurrence_orders(tr):
    """First order of the vector operators' recurrence form in the GEMM, or
    None: m0 with FP64 low orders, else 4 on a wave-resident grid (nlon =
    64 L; the first order tile keeps the two-segment form), else none."""
    if tr.vector_fp64_orders > 0:
        return tr.vector_fp64_orders
    nlon = tr.grid.nlon
    return 4 if nlon % 64 == 0 and nlon // 64 in (2, 3, 4, 6, 8, 12, 16, 24) else None


class RoundTrip:
    """The device buffers and the batch of one workload: the 121-field round
    trip, or with ``step`` the dycore's per-step set (243 grid outputs from
    the inverse, 201 grid fields into the forward)."""

    def __init__(self, specs, precision, torch, tune=True, split_acc=None, fp64_orders=None,
                 step=False):
        from arwen_spectral import SphericalHarmonicTransform, TransformBatch

        self.parts = []
        self.step = step
        jobs = []
        dev = torch.device("cuda", torch.cuda.current_device())
        for spec in specs:
            tr = SphericalHarmonicTransform.create(spec["T"], dealias_factor=1.0, precision=precision,
                                                   fp64_orders=fp64_orders)
            f = spec["fields"]
            c64 = lambda a: torch.from_numpy(a.astype(np.complex64)).to(dev)
            zeta = c64(f["vorticity"]).contiguous()
            div = c64(f["divergence"]).contiguous()
            M = zeta.shape[0]
            g = lambda n: torch.empty((M, n, *tr.grid.shape), dtype=torch.float32, device=dev)
            s = lambda n: torch.empty((M, n, *tr.spectral_shape), dtype=torch.complex64, device=dev)
            L = LEVELS
            if not step:
                scal = c64(np.concatenate([f["temperature"], f["lnps"]], axis=1)).contiguous()
                u, v, grid = g(L), g(L), g(L + 1)
                z2, d2, s2 = s(L), s(L), s(L + 1)
                pj = [("wind", tr, (zeta, div), (u, v)), ("inverse", tr, scal, grid),
                      ("vordiv", tr, (u, v), (z2, d2)), ("forward", tr, grid, s2)]
                part = dict(tr=tr, zeta=zeta, div=div, scal=scal, u=u, v=v, grid=grid,
                            z2=z2, d2=d2, s2=s2, spec=spec)
                part["pairs"] = {"vorticity": (z2, "vorticity"), "divergence": (d2, "divergence"),
                                 "temperature": (s2[:, :L], "temperature"),
                                 "lnps": (s2[:, L:], "lnps")}
            else:
                temp, lnq = c64(f["temperature"]).contiguous(), c64(f["lnq"]).contiguous()
                lnps = c64(f["lnps"]).contiguous()
                # geopotential on the grid: a resident input of the forward
                phig = tr.inverse(c64(f["geopotential"]))
                u, v, tg, tx, ty, q = g(L), g(L), g(L), g(L), g(L), g(L)
                lg, lx, ly = g(1), g(1), g(1)
                z2, d2, t2, q2, p2, l2 = s(L), s(L), s(L), s(L), s(L), s(1)
                pj = [("wind", tr, (zeta, div), (u, v)),
                      ("inverse_gradient", tr, temp, (tg, tx, ty)),
                      ("inverse", tr, lnq, q, {"exp": True}),      # q = exp(ln q) >= 0
                      ("inverse_gradient", tr, lnps, (lg, lx, ly)),
                      ("vordiv", tr, (u, v), (z2, d2)),
                      ("forward", tr, tg, t2), ("forward", tr, q, q2),
                      ("forward", tr, phig, p2), ("forward", tr, lg, l2)]
                part = dict(tr=tr, zeta=zeta, div=div, u=u, v=v, q=q, spec=spec)
                part["pairs"] = {"vorticity": (z2, "vorticity"), "divergence": (d2, "divergence"),
                                 "temperature": (t2, "temperature"), "lnps": (l2, "lnps"),
                                 "geopotential": (p2, "geopotential")}
            part["planes"] = [0, 0, 0, 0]  # GEMM planes: syn coef, syn waist, ana waist, ana coef
            part["tables"] = {"syn": set(), "ana": set()}
            part["rec_levels"] = 0  # vector levels in the recurrence form (both directions)
            for job in pj:
                nlev = job[2][0].shape[1] if isinstance(job[2], tuple) else job[2].shape[1]
                d, kinds = GEMM_TABLES[job[0]]
                if job[0] in ("wind", "vordiv") and recurrence_orders(tr) is not None:
                    kinds = "R"  # the basis to degree T+1 (the first tile's two tables: ~1%)
                    part["rec_levels"] += M * nlev
                part["tables"][d].update(kinds)
                for k, w in enumerate(GEMM_PLANES[job[0]]):
                    part["planes"][k] += M * nlev * w
            seq = lambda x: x if isinstance(x, tuple) else (x,)  # noqa: E731
            part["grid_out"] = sum(o.shape[0] * o.shape[1] for j in pj if GEMM_PLANES[j[0]][0]
                                   for o in seq(j[3]))
            part["grid_in"] = sum(i.shape[0] * i.shape[1] for j in pj if GEMM_PLANES[j[0]][2]
                                  for i in seq(j[2]))
            jobs += pj
            self.parts.append(part)
        self.batch = TransformBatch(jobs)
        if split_acc is not None:
            self.batch.set_config(split_acc=split_acc)
        # the autotuner picks each direction's column tile and grid by timing them
        self.config = self.batch.autotune(reps=3) if tune else self.batch.config()
        self.flops, self.fft_bytes, self.low_flops = self.batch.work()


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
        for name, (have, key) in part["pairs"].items():
            h = have[sel].cpu().numpy().astype(np.complex128)
            w = part["spec"]["fields"][key][sel]
            e = float(np.max(np.abs(h - w)) / np.max(np.abs(w)))
            key = f"T{part['spec']['T']}_{name}"
            out[key] = max(out.get(key, 0.0), e)
    return out


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or "unknown"


def _oracle_round_trip(o, zeta, div, scal):
    u, v = o.wind_from_vordiv(zeta, div)
    g = o.inverse(scal)
    z2, d2 = o.vordiv_from_wind(u, v)
    s2 = o.forward(g)
    return dict(u=u, v=v, grid=g, zeta=z2, div=d2, scal=s2)


def _time_oracle(o, fields, threads, warmups=2, iterations=5):
    """The reference CLI benchmark's protocol (spectral/cli.py:153-202): the
    first call timed on its own, then warmups, then the steady mean."""
    from threadpoolctl import threadpool_limits

    with threadpool_limits(threads):
        t0 = time.perf_counter()
        out = _oracle_round_trip(o, *fields)
        first = time.perf_counter() - t0
        for _ in range(warmups):
            _oracle_round_trip(o, *fields)
        times = []
        for _ in range(iterations):
            t0 = time.perf_counter()
            _oracle_round_trip(o, *fields)
            times.append(time.perf_counter() - t0)
    return first, float(np.mean(times)), out


def cpu_baseline(part):
    """FP64 oracle (the reference's per-order BLAS algorithm in numpy) on the
    host cores: the 121-field T383 round trip of member 0 on all cores (up to
    16), 2 warmups and 5 timed iterations; the same on one thread over a
    4-field sample (one level of vorticity, divergence, temperature + ln ps);
    and the T63 121-field round trip on all cores and on one thread.
    Returns the timing entry and the T383 oracle outputs."""
    sys.path.insert(0, str(ROOT))
    from oracle.sht import OracleTransform

    cores = min(len(os.sched_getaffinity(0)), 16)
    T = part["spec"]["T"]
    f = part["spec"]["fields"]
    full = (f["vorticity"][0], f["divergence"][0],
            np.concatenate([f["temperature"][0], f["lnps"][0]], axis=0))
    o = OracleTransform.create(T, dealias_factor=1.0)
    first, steady, out = _time_oracle(o, full, cores)
    sample = (full[0][:1], full[1][:1], np.concatenate([full[2][:1], full[2][LEVELS:]], axis=0))
    first1, steady1, _ = _time_oracle(o, sample, 1)
    t63 = member_fields(63, 1)
    o63 = OracleTransform.create(63, dealias_factor=1.0)
    f63 = (t63["vorticity"][0], t63["divergence"][0],
           np.concatenate([t63["temperature"][0], t63["lnps"][0]], axis=0))
    _, s63, _ = _time_oracle(o63, f63, cores)
    _, s63_1, _ = _time_oracle(o63, f63, 1)
    fields_total = 3 * LEVELS + 1
    entry = {
        "value": 1.0 / steady, "unit": "round trips/s", "cores": cores, "kind": "port",
        "cpu_model": cpu_model(),
        "sample": f"FP64 numpy oracle (the reference's per-order BLAS algorithm), T{T} "
                  f"{fields_total}-field round trip of member 0 (u, v from vorticity/divergence "
                  f"and back, temperature, ln ps) on {cores} threads: first call {first:.2f} s, "
                  f"2 warmups, steady mean of 5 iterations {steady:.2f} s",
        "first_call_s": first, "steady_s": steady, "warmups": 2, "iterations": 5,
        "single_thread": {"value": (4.0 / fields_total) / steady1, "unit": "round trips/s",
                          "cores": 1, "first_call_s": first1, "steady_s": steady1,
                          "sample": "4 of 121 fields (one level of vorticity, divergence, "
                                    "temperature + ln ps), 2 warmups + 5 iterations, scaled "
                                    "linearly to 121 fields (the oracle's per-call table "
                                    "expansion makes the scaled rate a lower bound)"},
        "c0_t63": {"value": 1.0 / s63, "unit": "round trips/s", "cores": cores,
                   "single_thread_value": 1.0 / s63_1,
                   "sample": "T63 64x128, 121 fields, 2 warmups + 5 iterations each"},
    }
    out["L"] = LEVELS
    return entry, out


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
    """Unique HBM bytes of the workload's two GEMM launches: each Legendre
    table a launch uses read once (P, dP/dphi, P m/cos: K_tri x nlat/2 folded
    elements, no padding; the basis to degree T+1 for the recurrence form:
    (T+2)(T+3)/2 x nlat/2; P and that one table are one array), and every
    coefficient plane (K_tri complex64) and waist plane (nlat x (T+1)
    complex64) the launch reads or writes, once."""
    total = 0.0
    for part in rt.parts:
        T, nlat = part["spec"]["T"], part["tr"].grid.nlat
        k_tri = (T + 1) * (T + 2) // 2
        k_ext = (T + 2) * (T + 3) // 2
        for kinds in part["tables"].values():
            ext = "R" in kinds
            n = len(set(kinds) - {"R", "P"}) + (1 if ext or "P" in kinds else 0)
            elems = (k_ext if ext else k_tri) * (nlat / 2)
            # P (or the extended basis) first, then the other kinds
            total += elems * TABLE_BYTES[precision] + (n - 1) * k_tri * (nlat / 2) * TABLE_BYTES[precision]
        coef = 8.0 * k_tri
        waist = 8.0 * nlat * (T + 1)
        sc, sw, aw, ac = part["planes"]
        total += sc * coef + sw * waist + aw * waist + ac * coef
    return total


def two_segment_equivalent(rt, precision, t, s, flops):
    """The GEMM launches' rate counted in the two-segment form's work: four
    plane contractions per vector level and direction above m0 (the
    reference's P m/cos and dP/dphi formulation), where the recurrence form
    runs two.  Not an MFMA utilisation (the MFMAs those FLOPs would take are
    not issued); the work-equivalent rate of the launches."""
    extra = 0.0
    for part in rt.parts:
        m0 = recurrence_orders(part["tr"])
        if m0 is None or not part["rec_levels"]:
            continue
        T, nlat = part["spec"]["T"], part["tr"].grid.nlat
        pcf = 4.0 * (nlat / 2) * (T + 1) * (T + 2) / 2
        pcf_low = 4.0 * (nlat / 2) * (m0 * (T + 1) - m0 * (m0 - 1) / 2)
        extra += 2.0 * part["rec_levels"] * (pcf - pcf_low)
    if extra == 0.0:
        return None
    f = flops + extra
    a = s * f / t / 1e12
    return {"flops": f, "achieved": a, "peak": MFMA_PEAK[precision], "unit": "TFLOP/s",
            "frac": a / MFMA_PEAK[precision]}


def roofline(rt, phases, precision, workload="t383"):
    """Roofline of the dominant kernel.  The Legendre GEMM (the grouped
    synthesis and analysis launches, timed by their own phase events) is
    priced on its MFMA roof: MFMA products per second of its algorithmic FLOPs
    against the dense peak of the instruction it issues; its HBM roof
    (algorithmic bytes per second) and the arithmetic intensity against the
    ridge point are reported beside it as ``other_roof``.  The vector jobs'
    orders m < m0 run in the FP64 low-order kernels, priced apart against the
    FP64 peak (``fp64_low_orders``); the FFT phases (all fold and unfold
    launches, which together can take longer than the two GEMM launches, each
    of them shorter) against HBM (``fft_phases``).  The MFMA figures count
    the plane contractions the launches run (two per vector level and
    direction in the recurrence form); ``two_segment_equivalent`` gives the
    same launches' rate in the two-segment form's work (four)."""
    gemm_ms = phases["synthesis_gemm"] + phases["analysis_gemm"]
    fft_ms = phases["inverse_fft"] + phases["forward_fft"]
    s = PRODUCTS[precision]
    # the grouped GEMMs' own FLOPs: the vector jobs' orders m < m0 run in the
    # FP64 low-order kernels, timed (and priced) apart
    flops = rt.flops - rt.low_flops
    low_ms = phases["synthesis_low"] + phases["analysis_low"]
    low = None
    if rt.low_flops > 0 and low_ms > 0:
        a = rt.low_flops / (low_ms * 1e-3) / 1e12
        low = {"kernel": "low_synthesis + low_analysis", "ms": low_ms,
               "flops": rt.low_flops, "achieved": a, "peak": FP64_PEAK_TFLOPS,
               "unit": "TFLOP/s", "frac": a / FP64_PEAK_TFLOPS}
    fft = None
    if fft_ms > 0:
        a = rt.fft_bytes / (fft_ms * 1e-3) / 1e9
        fft = {"kernel": "rfft_fold + irfft_unfold", "ms": fft_ms, "bytes": rt.fft_bytes,
               "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": a / HBM_PEAK_GBS}
    if gemm_ms > 0:
        t = gemm_ms * 1e-3
        alg_bytes = gemm_algorithmic_bytes(rt, precision)
        intensity = s * flops / alg_bytes
        ridge = MFMA_PEAK[precision] * 1e12 / (HBM_PEAK_GBS * 1e9)
        mfma = {"bound": "mfma", "achieved": s * flops / t / 1e12,
                "peak": MFMA_PEAK[precision], "unit": "TFLOP/s"}
        hbm = {"bound": "hbm", "achieved": alg_bytes / t / 1e9, "peak": HBM_PEAK_GBS,
               "unit": "GB/s"}
        for r in (mfma, hbm):
            r["frac"] = r["achieved"] / r["peak"]
        main, other = mfma, hbm
        return {**main, "kernel": "legendre_gemm (synthesis + analysis launches)",
                "traffic": gemm_traffic(rt.config, precision, workload),
                "traffic_unit": "HBM bytes per round trip (both launches)",
                "algorithmic_bytes": alg_bytes, "intensity_flop_per_byte": intensity,
                "ridge_flop_per_byte": ridge, "other_roof": other,
                "fp32_equivalent_tflops": flops / t / 1e12, "mfma_products_per_fma": s,
                "gemm_flops": flops, "fp64_low_orders": low, "fft_phases": fft,
                "two_segment_equivalent": two_segment_equivalent(rt, precision, t, s, flops)}
    achieved = rt.fft_bytes / (fft_ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": "rfft_fold + irfft_unfold", "achieved": achieved,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None}


SAMPLE_T, SAMPLE_FIELDS, SAMPLE_POINTS = 383, 40, 30000


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
    """Semi-implicit stage solve on spectral columns (solve_shifted, `value`):
    T383, 40 levels of divergence and theta plus ln ps, synthetic well-posed
    vertical operator, inputs resident on the device; HIP events on the
    launch stream.  Algorithmic bytes per call: the triangles of the 2L + 1
    input planes read once (complex64), the 2L + 1 output planes written
    once (whole (T+1)^2 squares: the outputs are fresh arrays, zero above
    the triangle), the per-degree inverses read once.  apply,
    linear_tendencies and half_apply are timed beside it on the same inputs
    (`other_operations`, each with its own algorithmic bytes: half_apply
    reads four per-degree matrices, tendencies none)."""
    import ctypes  # noqa: F401

    from arwen_spectral import SphericalHarmonicTransform
    from arwen_spectral.semi_implicit import (vertical_apply, vertical_half_apply,
                                              vertical_linear_tendencies, vertical_solve_shifted,
                                              wave_operator)

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
    cols = planes * (8.0 * k_tri + 8.0 * (T + 1) ** 2)  # triangles in, whole squares out
    alg = cols + 4.0 * (T + 1) * L * L
    achieved = alg / (ms * 1e-3) / 1e9
    # the other column operators on the same inputs: apply (with M and its
    # own inverse), linear_tendencies, half_apply (per-degree half maps)
    dt, alpha = 600.0, 0.6
    inv_a = f32(osi.helmholtz_inverse(modal, k2, dt, alpha))
    mm = f32(m)
    maps = [f32(x) for x in osi.symmetric_half_maps(modal, osi.x_modes(b, c), k2, dt, alpha)]
    ops = {
        "apply": (lambda: vertical_apply(t, d, th, ps, bm, cm, mm, inv_a, km, dt, alpha),
                  cols + 4.0 * (T + 1) * L * L),
        "linear_tendencies": (lambda: vertical_linear_tendencies(t, d, th, ps, bm, cm, km, 0.7),
                              cols),
        "half_apply": (lambda: vertical_half_apply(t, d, th, ps, bm, cm, *maps, km),
                       cols + 4.0 * (T + 1) * 2 * (L * L + (L + 1) ** 2)),
    }
    others = {}
    for name, (fn, nbytes) in ops.items():
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(args.steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        oms = e0.elapsed_time(e1) / args.steps
        gbs = nbytes / (oms * 1e-3) / 1e9
        others[name] = {"ms_per_call": oms, "algorithmic_bytes": nbytes, "achieved_gbs": gbs,
                        "frac": gbs / HBM_PEAK_GBS}
        print(f"[bench] vertical {name}: {oms:.3f} ms per call, {gbs / HBM_PEAK_GBS:.3f} of HBM",
              file=sys.stderr, flush=True)
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
        "roofline": {"bound": "hbm", "kernel": "vertical_rows_kernel", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "algorithmic_bytes": alg},
        "other_operations": others,
        "cpu_baseline": cpu}))


def ensemble_bench(args, torch):
    """Every member's grid fields (u, v, theta, qv, ln ps, psi, chi) of a
    32-member T383 L40 ensemble from one batch (EnsembleSynthesis), against
    the reference's member loop run on the engine (the same synthesis one
    member at a time); inputs resident, HIP events on the launch stream."""
    from arwen_spectral import EnsembleSynthesis, SphericalHarmonicTransform, TransformBatch

    T, R, L = 383, 32, LEVELS
    tr = SphericalHarmonicTransform.create(T, dealias_factor=1.0, precision=args.precision)
    ens = EnsembleSynthesis(tr, R, L)
    for name, buf in ens.inputs.items():
        lead = (R,) if name == "log_surface_pressure" else (R, L)
        scale = {"vorticity": 1e-5, "divergence": 2e-6, "theta": 10.0, "qv": 1e-3,
                 "log_surface_pressure": 1e-2}[name]
        buf.copy_(torch.from_numpy(red_spectrum(T, lead, 60 + len(name), scale)
                                   .astype(np.complex64)))
    ms_batch = _event_ms(torch, lambda: ens.run(graph=True), args.steps)
    # the member loop: the same syntheses as one batch per member, bound to
    # that member's slices of the same buffers, each replayed from its graph
    x, y = ens.inputs, ens.outputs
    singles = []
    for k in range(R):
        b = TransformBatch([
            ("wind", tr, (x["vorticity"][k], x["divergence"][k]), (y["u"][k], y["v"][k])),
            ("inverse", tr, x["theta"][k], y["theta"][k]),
            ("inverse", tr, x["qv"][k], y["qv"][k]),
            ("inverse", tr, x["log_surface_pressure"][k], y["lnps"][k]),
            ("inverse", tr, x["vorticity"][k], y["psi"][k], {"inverse_laplacian": True}),
            ("inverse", tr, x["divergence"][k], y["chi"][k], {"inverse_laplacian": True})])
        b.capture()
        singles.append(b)

    def loop():
        for b in singles:
            b.launch()
    ms_loop = _event_ms(torch, loop, max(1, args.steps // 4))
    fields = R * (6 * L + 1)
    print(json.dumps({
        "metric": f"T{T} ensemble grid fields/s ({R} members x (6 x {L} + 1) fields, one batch)",
        "value": fields / (ms_batch * 1e-3), "unit": "grid fields/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_batch,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": args.precision, "data": "synthetic: seeded red-spectrum member fields",
        "config": {"workload": f"EnsembleSynthesis T{T}, {R} members, {L} levels: u, v, theta, "
                               "qv, ln ps, psi, chi", "members": R, "levels": L},
        "member_loop_ms": ms_loop, "speedup_vs_member_loop": ms_loop / ms_batch,
        "roofline": None, "cpu_baseline": None}))


def workload_specs(name):
    """(specs, step): the truncations and members of a workload, and whether
    it is the per-step set rather than the 121-field round trip."""
    table = {
        "t383": ([dict(T=383, members=1)], False),
        "t383x32": ([dict(T=383, members=32)], False),
        "mixed": ([dict(T=767, members=4), dict(T=383, members=8), dict(T=127, members=16)],
                  False),
        "t63": ([dict(T=63, members=1)], False),
        "step": ([dict(T=383, members=1)], True),
        "stepx32": ([dict(T=383, members=32)], True),
    }
    if name not in table:
        raise SystemExit(f"unknown workload {name}; expected one of {sorted(table)} or sample, "
                         "vertical, ensemble")
    specs, step = table[name]
    return [dict(sp) for sp in specs], step


WORKLOAD_TEXT = {
    "t383": "T383 768x384 linear Gaussian grid, 121 fields per member (40 vorticity, 40 "
            "divergence, 40 temperature, ln ps): synthesis incl. u,v from vorticity/divergence, "
            "then analysis incl. vorticity/divergence from u,v",
    "step": "T383 768x384, the dycore's per-step transform set per member: inverse u,v from "
            "40 vorticity/divergence, 40 temperature with both gradients, 40 q = exp(ln q), "
            "ln ps with both gradients (243 grid outputs, 282 plane contractions); forward "
            "vorticity/divergence from 40 u,v, 40 temperature, 40 q, 40 geopotential, ln ps "
            "(201 grid fields, 281 plane contractions)",
}


def _event_ms(torch, fn, reps, stream=None):
    """Mean milliseconds of ``fn`` over ``reps`` calls, HIP events on the
    current stream after one untimed call."""
    fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def single_op_leg(torch, precision, reps=10):
    """Single calls at the dycore's chunk shapes (T383): a chunk of 6 fields x
    40 levels through inverse and forward, and a single ln ps plane through
    inverse, forward and gradient, each one grouped-GEMM launch per direction
    from the plan's cached workspaces (no allocation after the first call of a
    shape)."""
    from arwen_spectral import SphericalHarmonicTransform

    tr = SphericalHarmonicTransform.create(383, dealias_factor=1.0, precision=precision)
    dev = torch.device("cuda", torch.cuda.current_device())
    chunk = torch.from_numpy(red_spectrum(383, (6 * LEVELS,), 51, 10.0).astype(np.complex64)).to(dev)
    plane = chunk[:1].contiguous()
    gchunk = tr.inverse(chunk)
    gplane = gchunk[:1].contiguous()
    out = {}
    for name, fn, nf in (("inverse_6x40", lambda: tr.inverse(chunk), 240),
                         ("forward_6x40", lambda: tr.forward(gchunk), 240),
                         ("inverse_plane", lambda: tr.inverse(plane), 1),
                         ("forward_plane", lambda: tr.forward(gplane), 1),
                         ("gradient_plane", lambda: tr.gradient(plane), 1)):
        ms = _event_ms(torch, fn, reps)
        out[name] = {"ms_per_call": ms, "fields_per_s": nf / (ms * 1e-3)}
    return out


def gpu_t63(torch, precision, steps, warmup):
    """The T63 64x128 121-field round trip on the GPU (beside the oracle's
    T63 figure)."""
    specs, _ = workload_specs("t63")
    for sp in specs:
        sp["fields"] = member_fields(sp["T"], sp["members"])
    rt = RoundTrip(specs, precision, torch, True)
    sec = time_steps(rt, steps, warmup, torch)
    return {"value": 1.0 / sec, "unit": "round trips/s", "ms_per_step": sec * 1e3,
            "roundtrip_error": class_errors(rt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="t383")
    ap.add_argument("--precision", default="bf16x6")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the single-call and T63 legs")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--split-acc", type=int, default=0,
                    help="split modes: 1 = small products in their own accumulators; 0 (the "
                         "default) = one accumulator, which admits the 96-column GEMM tile and "
                         "keeps the FP32 path's error class")
    ap.add_argument("--fp64-orders", type=int, default=None,
                    help="vector operators: zonal orders below this run in FP64 end to end "
                         "(default: 0 for float32, 4 ceil((T+1)/192) for the split modes)")
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
    if args.workload == "ensemble":
        ensemble_bench(args, torch)
        return
    specs, step = workload_specs(args.workload)
    for sp in specs:
        sp["fields"] = member_fields(sp["T"], sp["members"], step=step)
    print(f"[bench] workload {args.workload}, precision {args.precision}", file=sys.stderr,
          flush=True)

    rt = RoundTrip(specs, args.precision, torch, not args.no_autotune, args.split_acc,
                   args.fp64_orders, step=step)
    sec = time_steps(rt, args.steps, args.warmup, torch)
    phases = rt.batch.profile(reps=3)
    errors = class_errors(rt)
    print(f"[bench] {args.precision}: {sec * 1e3:.3f} ms/step, phases {phases}, "
          f"gemm config {rt.config}", file=sys.stderr, flush=True)
    grid_out = sum(p["grid_out"] for p in rt.parts)
    grid_in = sum(p["grid_in"] for p in rt.parts)

    fp32 = None
    if not args.no_fp32 and args.precision != "float32":
        rt32 = RoundTrip(specs, "float32", torch, not args.no_autotune, step=step)
        sec32 = time_steps(rt32, args.steps, args.warmup, torch)
        ph32 = rt32.batch.profile(reps=3)
        fp32 = {"value": 1.0 / sec32, "ms_per_step": sec32 * 1e3, "phases_ms": ph32,
                "gemm_config": rt32.config,
                "roundtrip_error": class_errors(rt32),
                "roofline": roofline(rt32, ph32, "float32", args.workload)}
        print(f"[bench] float32: {sec32 * 1e3:.3f} ms/step", file=sys.stderr, flush=True)
        del rt32

    extras = None
    if not args.no_extras and args.workload == "t383":
        extras = {"single_calls": {args.precision: single_op_leg(torch, args.precision)},
                  "t63_gpu": {args.precision: gpu_t63(torch, args.precision, args.steps,
                                                      args.warmup)}}
        if not args.no_fp32 and args.precision != "float32":
            extras["single_calls"]["float32"] = single_op_leg(torch, "float32")
            extras["t63_gpu"]["float32"] = gpu_t63(torch, "float32", args.steps, args.warmup)
        print(f"[bench] extras {json.dumps(extras)}", file=sys.stderr, flush=True)

    cpu = None
    oracle_err = None
    if not args.no_cpu_baseline and not step:
        part = next((p for p in rt.parts if p["spec"]["T"] == 383), rt.parts[0])
        cpu, ref = cpu_baseline(part)
        oracle_err = oracle_errors(part, ref)
        print(f"[bench] cpu baseline {cpu['value']:.4f} round trips/s on {cpu['cores']} cores",
              file=sys.stderr, flush=True)

    members = sum(sp["members"] for sp in specs)
    unit = "steps/s" if step else "round trips/s"
    result = {
        "metric": METRIC if args.workload == "t383" else
                  f"{args.workload}: {unit} ({args.precision})",
        "value": 1.0 / sec,
        "unit": unit,
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
        "config": {"workload": WORKLOAD_TEXT.get(args.workload, args.workload),
                   "members": members, "grid_outputs_of_synthesis": grid_out,
                   "grid_inputs_of_analysis": grid_in, "precision": args.precision,
                   "vector_fp64_orders": {f"T{part['spec']['T']}": part["tr"].vector_fp64_orders
                                          for part in rt.parts},
                   "vector_form": "orders m < vector_fp64_orders in FP64, the others as one "
                                  "contraction of the basis per wind component (derivative "
                                  "recurrence)" if any(part["tr"].vector_fp64_orders
                                                       for part in rt.parts)
                                  else "orders m < 4 against the P m/cos and dP/dphi tables, the "
                                  "others as one contraction of the basis per wind component"
                                  if all(recurrence_orders(part["tr"]) for part in rt.parts)
                                  else "two table segments (P m/cos, dP/dphi)",
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
        "extras": extras,
        "cpu_baseline": cpu,
    }
    print(json.dumps(result))


if __name__ == "__main__":
    main()
