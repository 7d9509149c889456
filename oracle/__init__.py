"""FP64 CPU oracle for the spherical-harmonic transform path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product path
(``gpuwm-global_amd/``) imports, links or runs anything from this package:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use it, and there only as the checker or as the timed CPU
baseline, never as the thing measured or shipped.

``oracle.sht`` restates the reference's numpy transform
(``src/arwen_global/spectral/{grid,legendre,transform,vector}.py``, numpy
branches) in float64.  How each function is pinned is recorded in its
docstring and in ``tests/test_oracle_*.py``:

* ``gaussian_grid`` -- numpy's ``leggauss`` (the routine the reference
  calls), sum of weights = 2, exact quadrature to degree 2*nlat-1.
* ``legendre_columns`` / ``derivative_columns`` -- the reference's pinned
  sha256 table digests (numpy 2.2.6), the reference's scipy fixture values,
  live ``scipy.special.sph_harm_y``, closed forms, orthonormality, finite
  differences.
* ``analysis_rows`` -- closed-form harmonics analysed to 5e-15 and the
  round-trip / Parseval bounds of the reference's tests.
* the transform operations -- brute-force O(T^2 nlat nlon) sums at small T,
  solid-body rotation, degree-exact Laplacian, constant mode, vector round
  trip.

``oracle.sampling`` restates the reference's host point sampler
(``spectral/sampling.py``), pinned by live ``scipy.special.sph_harm_y`` at
random points and by the oracle's FFT synthesis, gradient and wind at the
Gaussian grid's points.
"""
