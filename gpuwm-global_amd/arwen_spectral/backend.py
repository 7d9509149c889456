"""Backend boundary (reference spectral/backend.py:14-71).

The reference switches its array module between numpy and CuPy.  This engine
has one backend, ``'hip'``: arrays are torch tensors on the current ROCm
device, and every transform runs in the HIP kernels of ``libarwen_sht.so``.
``precision`` names the Legendre GEMM arithmetic:

* ``'float32'``  FP32 MFMA on FP32 tables (the reference's float32 path);
* ``'bf16x6'``   split-BF16: tables and data as three BF16 pieces, the six
                 products down to 2^-16 relative accumulated in FP32 on MFMA;
* ``'bf16x3'``   split-BF16 hi+lo pairs, three products (~1e-5 relative).

Grid fields are float32 and spectral fields complex64 in every mode.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any

import numpy as np

from . import _native

PRECISIONS = {"float32": _native.FP32, "bf16x3": _native.BF16X3, "bf16x6": _native.BF16X6}


@dataclass(frozen=True)
class Backend:
    name: str
    xp: Any
    float_dtype: Any
    complex_dtype: Any
    precision: str = "float32"

    @property
    def precision_code(self) -> int:
        return PRECISIONS[self.precision]

    def device(self):
        return self.xp.device("cuda", self.xp.cuda.current_device())

    def asarray(self, value, *, dtype=None):
        torch = self.xp
        if isinstance(value, torch.Tensor):
            t = value
            if t.device.type != "cuda":
                t = t.to(self.device())
        else:
            arr = np.asarray(value)
            t = torch.from_numpy(np.ascontiguousarray(arr)).to(self.device())
        if dtype is not None and t.dtype != dtype:
            t = t.to(dtype)
        return t

    def to_numpy(self, value) -> np.ndarray:
        if isinstance(value, self.xp.Tensor):
            return value.detach().cpu().numpy()
        return np.asarray(value)

    def synchronize(self) -> None:
        self.xp.cuda.synchronize()

    def stream(self) -> int:
        return int(self.xp.cuda.current_stream().cuda_stream)


def get_backend(name: str = "hip", precision: str = "float32") -> Backend:
    key = str(name).strip().lower()
    prec = str(precision).strip().lower()
    if key != "hip":
        raise ValueError(
            f"backend must be 'hip', got {name!r}: this engine runs the transform "
            "in HIP kernels on the device only")
    if prec not in PRECISIONS:
        raise ValueError(
            f"precision must be one of {sorted(PRECISIONS)}, got {precision!r}")
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("backend='hip' needs a ROCm device; none is visible")
    _native.load()
    return Backend("hip", torch, torch.float32, torch.complex64, prec)
