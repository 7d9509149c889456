"""Build the HIP engine library in-tree for gfx950.

    python gpuwm-global_amd/build.py [--jobs N]

Compiles every source in ``csrc/`` with hipcc and links
``arwen_spectral/libarwen_sht.so``.  The Legendre table and point sampler
kernels are built with ``-ffp-contract=off`` so their FP64 recurrences keep
the reference's operand order without fused multiply-adds.  All device code is built with packed
FP32 instructions disabled.
"""
from __future__ import annotations

import argparse
import concurrent.futures as futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
INCLUDE = HERE.parent / "include"
OUT = HERE / "arwen_spectral" / "libarwen_sht.so"
OBJ = HERE / "_build"
ARCH = os.environ.get("ARWEN_SHT_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{CSRC}", f"-I{INCLUDE}",
]
EXTRA = {"sht_tables.hip": ["-ffp-contract=off"], "sht_sample.hip": ["-ffp-contract=off"]}
SOURCES = ["sht_tables.hip", "sht_fft.hip", "sht_gemm.hip", "sht_pack.hip", "sht_sample.hip",
           "sht_vertical.hip", "sht_lowm.hip", "sht_static.cpp", "sht_capi.cpp"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the engine needs the ROCm toolchain")


def compile_one(src: str) -> Path:
    obj = OBJ / (src + ".o")
    cmd = [hipcc(), *COMMON, *EXTRA.get(src, [])]
    if src.endswith(".cpp"):
        cmd += ["-x", "hip"]
    cmd += ["-c", str(CSRC / src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in res.stderr.splitlines() if noise not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{err}")
    if err.strip():
        print(err, file=sys.stderr)
    return obj


def build(jobs: int = 4) -> Path:
    OBJ.mkdir(exist_ok=True)
    with futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    print(build(ap.parse_args().jobs))
