"""The Fourier waist: the (*lead, nlat, T+1) complex buffer between the
longitude FFT and the Legendre contraction (reference
src/arwen_global/spectral/transform.py:68-94, 373-629).

On the way in it holds the zonal spectrum of each latitude row divided by
nlon and cut to m <= T; on the way out, the Legendre contraction's output
before the inverse FFT.  A band pipeline fills it (or drains it) a latitude
band at a time; the contraction always reads or writes it whole, so a banded
run returns the whole-field run's result.  The buffer is a torch tensor on
the engine's device; every fill and drain is one call into libarwen_sht.so.
"""
from __future__ import annotations

import numpy as np


def latitude_band_edges(nlat: int, bands: int) -> list[tuple[int, int]]:
    """Bands ``[floor(k nlat / B), floor((k+1) nlat / B))``, empty ones
    dropped, tiling ``[0, nlat)`` in ascending order (transform.py:68-94)."""
    n, b = int(nlat), int(bands)
    if n < 1:
        raise ValueError(f"nlat must be >= 1, got {nlat}")
    if b < 1:
        raise ValueError(f"bands must be >= 1, got {bands}")
    cuts = [(k * n) // b for k in range(b + 1)]
    return [(lo, hi) for lo, hi in zip(cuts, cuts[1:]) if hi > lo]


class _RowsFilled:
    """Which latitude rows a band fill has written (transform.py:373-414)."""

    def __init__(self, nlat: int):
        self._written = np.zeros(int(nlat), dtype=bool)

    def claim(self, r0: int, r1: int) -> None:
        if r1 <= r0 or r0 < 0 or r1 > self._written.size:
            raise ValueError(f"latitude band [{r0}, {r1}) is not a nonempty range of the "
                             f"{self._written.size} rows")
        if bool(self._written[r0:r1].any()):
            raise ValueError(f"latitude rows {r0}..{r1} are already filled: a second band "
                             "writing them would replace the first band's spectrum")
        self._written[r0:r1] = True

    def complete(self) -> bool:
        return bool(self._written.all())

    def require_complete(self) -> None:
        if not self.complete():
            missing = int((~self._written).sum())
            raise ValueError(f"{missing} of {self._written.size} latitude rows of the Fourier "
                             "waist were never filled; contracting it would read unset memory")


class FourierWaist:
    """A named (*lead, nlat, T+1) waist with a band schedule and a lifetime:
    :meth:`take` hands the buffer to its last reader, after which the waist
    is consumed (transform.py:417-629)."""

    def __init__(self, transform, values, edges):
        self.transform = transform
        self._values = values
        self._open = False
        self._filled = None
        self.edges = list(edges)
        self.lead = tuple(values.shape[:-2])
        self.shape = tuple(values.shape)
        self.nbytes = int(values.numel() * values.element_size())

    @classmethod
    def open(cls, transform, lead, edges) -> "FourierWaist":
        """An empty waist for :meth:`fill_band` (transform.py:452-481)."""
        torch = transform.backend.xp
        lead = tuple(int(v) for v in lead)
        values = torch.empty((*lead, transform.grid.nlat, transform.truncation + 1),
                             dtype=transform.backend.complex_dtype,
                             device=transform.backend.device())
        waist = cls(transform, values, edges)
        waist._open = True
        waist._filled = _RowsFilled(transform.grid.nlat)
        return waist

    def fill_band(self, r0: int, r1: int, rows) -> None:
        """Write the zonal spectrum of latitude rows ``r0 .. r1`` (grid rows
        ``(*lead, r1 - r0, nlon)``) into the waist (transform.py:483-525)."""
        if not self._open:
            raise RuntimeError("this FourierWaist was not opened for band fills: a waist built "
                               "from a whole field is already full")
        t = self.transform
        r0, r1 = int(r0), int(r1)
        self._filled.claim(r0, r1)
        f = t.backend.asarray(rows, dtype=t.backend.float_dtype).contiguous()
        want = (*self.lead, r1 - r0, t.grid.nlon)
        if tuple(f.shape) != want:
            raise ValueError(f"band rows {tuple(f.shape)} are not the waist's {want}")
        t._waist_from_rows(f, self.values, r0, r1 - r0)

    def close(self) -> "FourierWaist":
        """Refuse a waist with an unwritten row, then hand it back
        (transform.py:527-566; one device, so no rows arrive from others)."""
        if self._open:
            self._filled.require_complete()
            self._open = False
        return self

    @property
    def bands(self) -> int:
        return len(self.edges)

    @property
    def values(self):
        if self._values is None:
            raise RuntimeError("this FourierWaist was consumed: its buffer went to a drain or "
                               "a contraction; build a new waist instead")
        return self._values

    def take(self):
        """Hand the buffer over and forget it (transform.py:588-591)."""
        values = self.values
        self._values = None
        return values

    def band(self, r0: int, r1: int):
        """The waist rows of one latitude band, a view."""
        return self.values[..., int(r0):int(r1), :]

    def drain_scratch(self, rows: int):
        """A zero ``(*lead, rows, nlon//2+1)`` half spectrum for one band
        (transform.py:596-629); the engine's drains do not need it, callers
        that build their own half spectra may."""
        t = self.transform
        torch = t.backend.xp
        return torch.zeros((*self.lead, int(rows), t.grid.nlon // 2 + 1),
                           dtype=t.backend.complex_dtype, device=t.backend.device())
