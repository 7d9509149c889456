"""Grouped transform batches: one persistent Legendre GEMM launch per direction.

A ``TransformBatch`` takes an ordered list of jobs -- synthesis-type jobs
(``inverse``, ``zonal``, ``meridional``, ``gradient``, ``wind``) and
analysis-type jobs (``forward``, ``vordiv``) on any mix of transforms (one
precision) -- and runs them as

    pack -> one grouped synthesis GEMM -> inverse FFTs
    forward FFTs -> one grouped analysis GEMM -> unpack

so all fields, levels, ensemble members and truncations of the batch share
one launch per direction.  An analysis job may read a synthesis job's output
of the same batch (a round trip).  ``capture()`` records the launches in a
hipGraph; ``launch()`` replays it.  The reference has no counterpart: it
transforms fields in chunks and ensemble members one at a time
(reference da/ensemble.py:437-455).
"""
from __future__ import annotations

import ctypes

from . import _native
from .transform import SphericalHarmonicTransform

OPS = {
    "inverse": (_native.OP_SYNTHESIS, 1, 1),
    "inverse_gradient": (_native.OP_SYNTHESIS_GRADIENT, 1, 3),
    "zonal": (_native.OP_ZONAL_DERIVATIVE, 1, 1),
    "meridional": (_native.OP_MERIDIONAL_DERIVATIVE, 1, 1),
    "gradient": (_native.OP_GRADIENT, 1, 2),
    "wind": (_native.OP_UV_FROM_VORDIV, 2, 2),
    "forward": (_native.OP_ANALYSIS, 1, 1),
    "vordiv": (_native.OP_VORDIV_FROM_UV, 2, 2),
}
_SPECTRAL_IN = {"inverse", "inverse_gradient", "zonal", "meridional", "gradient", "wind"}
#: per-job options: "exp" (inverse: write exp of the grid),
#: "inverse_laplacian" (inverse: synthesize the inverse Laplacian of the
#: input, the streamfunction or velocity potential of a vorticity or
#: divergence), "grid_offset" (forward: float32 (count,) subtracted from each
#: field's grid first), "coeff_offset" (forward, inverse, inverse_gradient:
#: complex64 (count,) added to each field's (0, 0) coefficient, after the
#: analysis or before the synthesis)
_OPTION_OPS = {"exp": {"inverse"}, "inverse_laplacian": {"inverse"}, "grid_offset": {"forward"},
               "coeff_offset": {"forward", "inverse", "inverse_gradient"}}


class TransformBatch:
    """``jobs``: list of ``(op, transform, inputs, outputs)`` or ``(op, transform,
    inputs, outputs, options)`` with torch tensors; ``inverse_gradient``
    writes (grid, east, north) from one spectral input.

    Inputs and outputs are caller-owned device tensors in the boundary
    layouts; they must stay alive and keep their storage for the batch's
    lifetime (the work list holds their addresses).
    """

    def __init__(self, jobs):
        self._lib = _native.load()
        self._keep = []
        arr = (_native.Job * len(jobs))()
        precision = None
        for k, job in enumerate(jobs):
            op, tr, ins, outs = job[:4]
            opts = dict(job[4]) if len(job) > 4 and job[4] else {}
            if op not in OPS:
                raise ValueError(f"unknown op {op!r}; expected one of {sorted(OPS)}")
            if not isinstance(tr, SphericalHarmonicTransform):
                raise ValueError("each job needs a SphericalHarmonicTransform")
            code, nin, nout = OPS[op]
            ins = tuple(ins) if isinstance(ins, (tuple, list)) else (ins,)
            outs = tuple(outs) if isinstance(outs, (tuple, list)) else (outs,)
            if len(ins) != nin or len(outs) != nout:
                raise ValueError(f"{op} takes {nin} input(s) and {nout} output(s)")
            want_in = tr.spectral_shape if op in _SPECTRAL_IN else tr.grid.shape
            want_out = tr.grid.shape if op in _SPECTRAL_IN else tr.spectral_shape
            lead = tuple(ins[0].shape[:-2])
            for t in ins:
                if tuple(t.shape[-2:]) != want_in or tuple(t.shape[:-2]) != lead:
                    raise ValueError(f"{op} input has shape {tuple(t.shape)}")
                if not t.is_contiguous():
                    raise ValueError(f"{op} inputs must be contiguous")
            for t in outs:
                if tuple(t.shape[-2:]) != want_out or tuple(t.shape[:-2]) != lead:
                    raise ValueError(f"{op} output has shape {tuple(t.shape)}")
                if not t.is_contiguous():
                    raise ValueError(f"{op} outputs must be contiguous")
            if precision is None:
                precision = tr.backend.precision
            elif precision != tr.backend.precision:
                raise ValueError("all transforms of a batch must share one precision")
            count = SphericalHarmonicTransform._count(ins[0])
            for name in opts:
                if name not in _OPTION_OPS or op not in _OPTION_OPS[name]:
                    raise ValueError(f"option {name!r} does not apply to {op}")
            flags = (_native.JOB_EXP if opts.get("exp") else 0) | \
                (_native.JOB_INVLAP if opts.get("inverse_laplacian") else 0)
            if opts.get("inverse_laplacian") and opts.get("coeff_offset") is not None:
                raise ValueError("inverse_laplacian takes no coeff_offset")
            offsets = []
            for name, dtype in (("grid_offset", "float32"), ("coeff_offset", "complex64")):
                t = opts.get(name)
                if t is not None:
                    if str(t.dtype) != f"torch.{dtype}" or t.numel() != count \
                            or not t.is_contiguous() or t.device.type != "cuda":
                        raise ValueError(f"{name} must be a contiguous {dtype} device tensor of "
                                         f"{count} entries")
                offsets.append(t.data_ptr() if t is not None else None)
            ptr = lambda seq, i: seq[i].data_ptr() if i < len(seq) else None
            arr[k] = _native.Job(code, tr.plan_handle.value, count, ptr(ins, 0), ptr(ins, 1),
                                 ptr(outs, 0), ptr(outs, 1), ptr(outs, 2), flags, *offsets)
            self._keep.append((tr, ins, outs, opts))
        handle = ctypes.c_void_p()
        _native.check(self._lib.arwen_sht_batch_create(arr, len(jobs), ctypes.byref(handle)))
        self._handle = handle
        self._backend = jobs[0][1].backend

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h:
            try:
                self._lib.arwen_sht_batch_destroy(h)
            except Exception:
                pass
            self._handle = None

    def run(self):
        _native.check(self._lib.arwen_sht_batch_run(self._handle, self._backend.stream()))

    def capture(self):
        _native.check(self._lib.arwen_sht_batch_capture(self._handle, self._backend.stream()))

    def launch(self):
        _native.check(self._lib.arwen_sht_batch_launch(self._handle, self._backend.stream()))

    def config(self) -> dict:
        """The grouped-GEMM schedule: ``cols`` and ``grid`` per direction
        (synthesis, analysis) and ``split_acc``."""
        c = _native.GemmConfig()
        _native.check(self._lib.arwen_sht_batch_get_config(self._handle, ctypes.byref(c)))
        return c.as_dict()

    def set_config(self, cols=None, grid=None, split_acc=None):
        """Change the schedule; ``cols``/``grid`` are (synthesis, analysis) pairs."""
        c = _native.GemmConfig()
        _native.check(self._lib.arwen_sht_batch_get_config(self._handle, ctypes.byref(c)))
        if cols is not None:
            c.cols[0], c.cols[1] = (int(v) for v in cols)
        if grid is not None:
            c.grid[0], c.grid[1] = (int(v) for v in grid)
        if split_acc is not None:
            c.split_acc = int(split_acc)
        _native.check(self._lib.arwen_sht_batch_set_config(self._handle, ctypes.byref(c)))

    def autotune(self, reps: int = 3) -> dict:
        """Time every column tile and grid per direction on the current stream
        and keep the fastest (synchronizes; overwrites intermediate buffers)."""
        c = _native.GemmConfig()
        _native.check(self._lib.arwen_sht_batch_autotune(self._handle, self._backend.stream(),
                                                         int(reps), ctypes.byref(c)))
        return c.as_dict()

    PHASES = ("pack", "synthesis_gemm", "synthesis_low", "inverse_fft", "forward_fft",
              "analysis_gemm", "analysis_low", "unpack")

    def profile(self, reps: int = 5) -> dict:
        """Mean milliseconds per phase over ``reps`` runs, HIP events on the
        current stream (synchronizes)."""
        import numpy as np

        ms = np.zeros(len(self.PHASES), dtype=np.float64)
        _native.check(self._lib.arwen_sht_batch_profile(self._handle, self._backend.stream(),
                                                        int(reps), _native.dptr(ms)))
        return dict(zip(self.PHASES, ms.tolist()))

    def work(self) -> tuple[float, float, float]:
        """(Legendre FLOPs, FFT bytes, the FP64 low-order kernels' share of
        the Legendre FLOPs) per run, algorithmic."""
        f = ctypes.c_double()
        b = ctypes.c_double()
        low = ctypes.c_double()
        _native.check(self._lib.arwen_sht_batch_work(self._handle, ctypes.byref(f), ctypes.byref(b),
                                                     ctypes.byref(low)))
        return f.value, b.value, low.value
