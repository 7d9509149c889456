"""ctypes binding of ``libarwen_sht.so`` (the C ABI in ``include/arwen_sht.h``).

The library is built in-tree by ``gpuwm-global_amd/build.py``.  There is no
fallback: if the library is missing or cannot be loaded every entry point
raises, so no result can silently come from anywhere but the HIP kernels.
"""
from __future__ import annotations

import ctypes
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libarwen_sht.so"

OK, EINVAL, EHIP, ENOMEM = 0, 1, 2, 3
FP32, BF16X3, BF16X6 = 1, 2, 3
OP_SYNTHESIS = 1
OP_ZONAL_DERIVATIVE = 2
OP_MERIDIONAL_DERIVATIVE = 3
OP_GRADIENT = 4
OP_UV_FROM_VORDIV = 5
OP_ANALYSIS = 6
OP_VORDIV_FROM_UV = 7
OP_SYNTHESIS_TO_WAIST = 8
OP_ZONAL_TO_WAIST = 9
OP_MERIDIONAL_TO_WAIST = 10
OP_ANALYSIS_FROM_WAIST = 11
OP_VORDIV_FROM_WAIST = 12
OP_SYNTHESIS_GRADIENT = 13
OP_DIVERGENCE_FROM_WAIST = 14
JOB_EXP = 1
JOB_INVLAP = 2
# kind argument of arwen_sht_synthesis_to_waist
WAIST_BASIS, WAIST_ZONAL, WAIST_DERIVATIVE = 0, 1, 2

SPECTRAL_PROJECT, SPECTRAL_LAPLACIAN, SPECTRAL_INVERSE_LAPLACIAN = 0, 1, 2
#: ARWEN_SHT_SAMPLE_STATIC_BYTES: the static sampler's per-GPU workspace array
SAMPLE_STATIC_BYTES = 64 << 20

#: Every symbol the header declares; the CPU suite checks the library exports them.
EXPORTS = (
    "arwen_sht_plan_create", "arwen_sht_plan_destroy", "arwen_sht_plan_grid",
    "arwen_sht_gauss_nodes", "arwen_sht_plan_table_bytes", "arwen_sht_plan_fp64_orders",
    "arwen_sht_plan_table_f64",
    "arwen_sht_analysis", "arwen_sht_synthesis", "arwen_sht_zonal_derivative",
    "arwen_sht_meridional_derivative", "arwen_sht_gradient", "arwen_sht_synthesis_gradient",
    "arwen_sht_analysis_about", "arwen_sht_synthesis_about", "arwen_sht_uv_from_vordiv",
    "arwen_sht_vordiv_from_uv", "arwen_sht_synthesis_to_waist", "arwen_sht_analysis_from_waist",
    "arwen_sht_vordiv_from_waist", "arwen_sht_divergence_from_waist", "arwen_sht_waist_from_rows", "arwen_sht_rows_from_waist",
    "arwen_sht_spectral_op", "arwen_sht_sample_workspace_bytes", "arwen_sht_sample",
    "arwen_sht_sample_static",
    "arwen_sht_vertical_solve", "arwen_sht_vertical_tendencies", "arwen_sht_vertical_half_apply",
    "arwen_sht_batch_create",
    "arwen_sht_batch_destroy", "arwen_sht_batch_run", "arwen_sht_batch_capture",
    "arwen_sht_batch_launch", "arwen_sht_batch_get_config", "arwen_sht_batch_set_config",
    "arwen_sht_batch_autotune", "arwen_sht_batch_work", "arwen_sht_batch_profile",
    "arwen_sht_last_error",
)


class PlanDesc(ctypes.Structure):
    _fields_ = [
        ("truncation", ctypes.c_int),
        ("nlat", ctypes.c_int),
        ("nlon", ctypes.c_int),
        ("radius_m", ctypes.c_double),
        ("precision", ctypes.c_int),
        ("sin_lat", ctypes.POINTER(ctypes.c_double)),
        ("weights", ctypes.POINTER(ctypes.c_double)),
        ("fp64_orders", ctypes.c_int),
    ]


class GemmConfig(ctypes.Structure):
    """``arwen_sht_gemm_config``: per direction (0 synthesis, 1 analysis)."""

    _fields_ = [
        ("cols", ctypes.c_int * 2),
        ("grid", ctypes.c_int * 2),
        ("split_acc", ctypes.c_int),
    ]

    def as_dict(self) -> dict:
        return {"cols": list(self.cols), "grid": list(self.grid), "split_acc": self.split_acc}


class Job(ctypes.Structure):
    _fields_ = [
        ("op", ctypes.c_int),
        ("plan", ctypes.c_void_p),
        ("count", ctypes.c_int),
        ("in0", ctypes.c_void_p),
        ("in1", ctypes.c_void_p),
        ("out0", ctypes.c_void_p),
        ("out1", ctypes.c_void_p),
        ("out2", ctypes.c_void_p),
        ("flags", ctypes.c_int),
        ("grid_offset", ctypes.c_void_p),
        ("coeff_offset", ctypes.c_void_p),
    ]


_lock = threading.Lock()
_lib = None


def load():
    """The loaded library; raises if it is not built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        # torch's bundled HIP runtime carries the same SONAME as the one the
        # library links (libamdhip64.so.7); loading torch first makes the
        # dynamic linker bind the library to that one runtime instead of
        # mapping a second HIP runtime into the process.
        import torch  # noqa: F401

        if not LIB_PATH.exists():
            raise RuntimeError(
                f"the HIP engine library {LIB_PATH} is not built; run "
                "`python gpuwm-global_amd/build.py` (or __graft_entry__.build())"
            )
        lib = ctypes.CDLL(str(LIB_PATH))
        P = ctypes.c_void_p
        I = ctypes.c_int
        D = ctypes.POINTER(ctypes.c_double)
        F = ctypes.c_float
        sigs = {
            "arwen_sht_plan_create": (I, [ctypes.POINTER(PlanDesc), ctypes.POINTER(P)]),
            "arwen_sht_plan_destroy": (I, [P]),
            "arwen_sht_plan_grid": (I, [P, D, D]),
            "arwen_sht_gauss_nodes": (I, [I, D, D]),
            "arwen_sht_plan_table_bytes": (ctypes.c_size_t, [P]),
            "arwen_sht_plan_fp64_orders": (I, [P]),
            "arwen_sht_plan_table_f64": (I, [P, I, D]),
            "arwen_sht_analysis": (I, [P, P, P, I, P]),
            "arwen_sht_synthesis": (I, [P, P, P, I, P]),
            "arwen_sht_zonal_derivative": (I, [P, P, P, I, P]),
            "arwen_sht_meridional_derivative": (I, [P, P, P, I, P]),
            "arwen_sht_gradient": (I, [P, P, P, P, I, P]),
            "arwen_sht_synthesis_gradient": (I, [P, P, P, P, P, I, P]),
            "arwen_sht_analysis_about": (I, [P, P, P, I, P, P, P]),
            "arwen_sht_synthesis_about": (I, [P, P, P, I, P, I, P]),
            "arwen_sht_uv_from_vordiv": (I, [P, P, P, P, P, I, P]),
            "arwen_sht_vordiv_from_uv": (I, [P, P, P, P, P, I, P]),
            "arwen_sht_synthesis_to_waist": (I, [P, P, P, I, I, P]),
            "arwen_sht_analysis_from_waist": (I, [P, P, P, I, P]),
            "arwen_sht_vordiv_from_waist": (I, [P, P, P, P, P, I, P]),
            "arwen_sht_divergence_from_waist": (I, [P, P, P, P, I, P]),
            "arwen_sht_waist_from_rows": (I, [P, P, P, I, I, I, P]),
            "arwen_sht_rows_from_waist": (I, [P, P, P, I, I, I, P]),
            "arwen_sht_spectral_op": (I, [P, I, P, P, I, P]),
            "arwen_sht_sample_workspace_bytes": (ctypes.c_size_t, [P, I, I, I]),
            "arwen_sht_sample": (I, [P, P, P, I, P, P, P, I, I, I, P, ctypes.c_size_t, P, P, P]),
            "arwen_sht_sample_static": (I, [P, P, P, I, P, P, P, I, I, I, P, P, P]),
            "arwen_sht_vertical_solve": (I, [P, I, P, P, P, P, P, P, P, P, F, F, F, F, F,
                                             P, P, P, P]),
            "arwen_sht_vertical_tendencies": (I, [P, I, P, P, P, P, P, P, F, P, P, P, P]),
            "arwen_sht_vertical_half_apply": (I, [P, I, P, P, P, P, P, P, P, P, P, P, P, P, P,
                                                  P]),
            "arwen_sht_batch_create": (I, [ctypes.POINTER(Job), I, ctypes.POINTER(P)]),
            "arwen_sht_batch_destroy": (I, [P]),
            "arwen_sht_batch_run": (I, [P, P]),
            "arwen_sht_batch_capture": (I, [P, P]),
            "arwen_sht_batch_launch": (I, [P, P]),
            "arwen_sht_batch_get_config": (I, [P, ctypes.POINTER(GemmConfig)]),
            "arwen_sht_batch_set_config": (I, [P, ctypes.POINTER(GemmConfig)]),
            "arwen_sht_batch_autotune": (I, [P, P, I, ctypes.POINTER(GemmConfig)]),
            "arwen_sht_batch_work": (I, [P, D, D, D]),
            "arwen_sht_batch_profile": (I, [P, P, I, D]),
            "arwen_sht_last_error": (ctypes.c_char_p, []),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def check(rc: int) -> None:
    """Re-raise a C status as the reference's exception types."""
    if rc == OK:
        return
    msg = load().arwen_sht_last_error().decode(errors="replace")
    if rc == EINVAL:
        raise ValueError(msg)
    if rc == ENOMEM:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def dptr(array) -> ctypes.POINTER(ctypes.c_double):
    return array.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
