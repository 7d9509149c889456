// Spectral fields evaluated at arbitrary points (reference
// spectral/sampling.py:382-515: sample_scalar, sample_gradient, sample_wind).
//
// One lane per point, one block per (256 points, tile of FT fields).  The lane
// walks the whole triangle order by order: the normalized associated-Legendre
// recurrence of its point in FP64 (reference legendre.py:86-146, the
// reference's operand order in every expression; this file is compiled with
// -ffp-contract=off and the contractions below are explicit fma), and for
// each (n, m) the tile's coefficients, which every lane of the block reads
// at the same address, multiply-add into per-order amplitudes held in
// registers.  Each order's amplitude is turned by the zonal phase
// e^{i m lambda} and added into FP64 outputs.  No basis cube is stored: the
// reference's device path materialises a (2K x points) basis per point chunk
// for a GEMM, which on MI355X buys nothing (the FP64 and FP32 matrix peaks
// equal the vector ones) and costs 8 bytes per basis value of HBM traffic.
//
// Precision: F32 = false accumulates in FP64; F32 = true rounds each basis
// value to fp32 and accumulates an order's terms in fp32 before the FP64 sum
// (the reference's "float32" sampler, sampling.py:30-44).
#include "sht_kernels.h"

#include <math.h>

#include <type_traits>

namespace sht {

// Packed triangle index of (n, m), order-major (reference sampling.py:72-80).
__device__ __forceinline__ long long tri_index(int T, int n, int m) {
  return (long long)m * (T + 1) - ((long long)m * (m - 1)) / 2 + (n - m);
}

// (count, T+1, T+1) complex64 [n, m] -> (K, ldc) complex, fields fastest:
// complex128 for the float64 sampler (the FMAs then take the coefficient
// straight from scalar registers), complex64 for the float32 one; columns
// past the live fields are zero.  Winds: column 2l is psi and 2l + 1 chi of
// level l, from vorticity / divergence by the inverse Laplacian.
__global__ void sample_pack_kernel(SamplePackArgs a) {
  const int T1 = a.T + 1;
  const long long total = (long long)a.ldc * T1 * T1;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int m = (int)(e % T1);
    const int n = (int)((e / T1) % T1);
    const int col = (int)(e / ((long long)T1 * T1));
    if (m > n) continue;
    float2 v = make_float2(0.f, 0.f);
    if (a.wind) {
      // the inverse Laplacian in the arithmetic of the spectral operator
      // kernel (c / lambda_n in fp32, lambda_n = -n(n+1)/a^2 rounded once),
      // so psi and chi are inverse_laplacian's bits
      const int level = col >> 1;
      if (level < a.count && n > 0) {
        const float2* src = (col & 1) ? a.in1 : a.in0;
        const float2 c = src[((size_t)level * T1 + n) * T1 + m];
        const float eig = (float)(-((double)n * ((double)n + 1.0)) / (a.radius * a.radius));
        v = make_float2(c.x / eig, c.y / eig);
      }
    } else if (col < a.count) {
      v = a.in0[((size_t)col * T1 + n) * T1 + m];
    }
    const long long o = tri_index(a.T, n, m) * a.ldc + col;
    if (a.f64) static_cast<double2*>(a.out)[o] = make_double2(v.x, v.y);
    else static_cast<float2*>(a.out)[o] = v;
  }
}

template <bool F32>
struct SampleAcc {
  using T = typename std::conditional<F32, float, double>::type;
  using C = typename std::conditional<F32, float2, double2>::type;
};

// KIND 0: the field; 1: east and north derivatives per metre; 2: u and v
// from column pairs (psi, chi).
template <int FT, int KIND, bool F32>
__global__ void __launch_bounds__(256) sample_kernel(
    const typename SampleAcc<F32>::C* __restrict__ coef, const double* __restrict__ rec,
    const double* __restrict__ diag, const double* __restrict__ sin_lat,
    const double* __restrict__ cos_lat, const double* __restrict__ lon, int T, int npts,
    int ldc, int nout, double radius, double* __restrict__ out0, double* __restrict__ out1) {
  using Acc = typename SampleAcc<F32>::T;
  constexpr bool GRAD = KIND != 0;
  const int p = blockIdx.x * 256 + threadIdx.x;
  const bool live = p < npts;
  const int pp = live ? p : npts - 1;  // idle lanes evaluate a real point, store nothing
  const int c0 = blockIdx.y * FT;
  const long long K = (long long)(T + 1) * (T + 2) / 2;
  const double* rec_a = rec;
  const double* rec_b = rec + K;
  const double* rec_c = rec + 2 * K;
  const double* diag_f = diag;
  const double* diag_g = diag + (T + 1);
  const typename SampleAcc<F32>::C* cf = coef + c0;

  const double x = sin_lat[pp];
  const double lam = lon[pp];
  const double root = sqrt(fmax(0.0, 1.0 - x * x));
  const double denom = x * x - 1.0;
  const double pi = 3.14159265358979323846;
  double dg = 1.0 / sqrt(4.0 * pi);

  double o0[FT], o1[FT];
#pragma unroll
  for (int f = 0; f < FT; ++f) o0[f] = o1[f] = 0.0;

  long long k = 0;
  for (int m = 0; m <= T; ++m) {
    if (m > 0) dg = diag_f[m] * root * dg;
    Acc ar[FT], ai[FT], dr[FT], di[FT];
#pragma unroll
    for (int f = 0; f < FT; ++f) ar[f] = ai[f] = dr[f] = di[f] = Acc(0);
    auto add = [&](long long kk, double P, double dP) {
      const auto* c = cf + kk * ldc;
#pragma unroll
      for (int f = 0; f < FT; ++f) {
        const auto v = c[f];
        if constexpr (F32) {
          ar[f] = fmaf(v.x, (float)P, ar[f]);
          ai[f] = fmaf(v.y, (float)P, ai[f]);
          if constexpr (GRAD) {
            dr[f] = fmaf(v.x, (float)dP, dr[f]);
            di[f] = fmaf(v.y, (float)dP, di[f]);
          }
        } else {
          ar[f] = fma(v.x, P, ar[f]);
          ai[f] = fma(v.y, P, ai[f]);
          if constexpr (GRAD) {
            dr[f] = fma(v.x, dP, dr[f]);
            di[f] = fma(v.y, dP, di[f]);
          }
        }
      }
    };
    // n = m
    double prev = dg, prev2 = 0.0;
    double dP = 0.0;
    if constexpr (GRAD) dP = (m >= 1) ? root * ((double)m * x * dg) / denom : 0.0;
    add(k, dg, dP);
    ++k;
    if (m < T) {
      // n = m + 1
      const double p1 = diag_g[m] * x * dg;
      if constexpr (GRAD) dP = root * ((double)(m + 1) * x * p1 - rec_c[k] * prev) / denom;
      add(k, p1, dP);
      ++k;
      prev2 = prev;
      prev = p1;
      for (int n = m + 2; n <= T; ++n) {
        const double pn = rec_a[k] * (x * prev - rec_b[k] * prev2);
        if constexpr (GRAD) dP = root * ((double)n * x * pn - rec_c[k] * prev) / denom;
        add(k, pn, dP);
        ++k;
        prev2 = prev;
        prev = pn;
      }
    }
    double sm, cm;
    sincos((double)m * lam, &sm, &cm);
    const double mult = m == 0 ? 1.0 : 2.0;
#pragma unroll
    for (int f = 0; f < FT; ++f) {
      const double re = (double)ar[f], im = (double)ai[f];
      if constexpr (KIND == 0) {
        o0[f] += mult * (re * cm - im * sm);
      } else {
        // Re(i m A e^{i m lambda}) and Re(dA e^{i m lambda})
        o0[f] += mult * (-(double)m * (re * sm + im * cm));
        o1[f] += mult * ((double)dr[f] * cm - (double)di[f] * sm);
      }
    }
  }
  if (!live) return;
  if constexpr (KIND == 0) {
#pragma unroll
    for (int f = 0; f < FT; ++f)
      if (c0 + f < nout) out0[(size_t)(c0 + f) * npts + p] = o0[f];
  } else if constexpr (KIND == 1) {
    const double east_scale = radius * cos_lat[p];
#pragma unroll
    for (int f = 0; f < FT; ++f)
      if (c0 + f < nout) {
        out0[(size_t)(c0 + f) * npts + p] = o0[f] / east_scale;
        out1[(size_t)(c0 + f) * npts + p] = o1[f] / radius;
      }
  } else {
    // columns 2l, 2l + 1 = psi, chi of level l: u = -north(psi) + east(chi),
    // v = east(psi) + north(chi) (reference sampling.py:481-515)
    static_assert(FT % 2 == 0, "wind tiles hold whole (psi, chi) pairs");
    const double east_scale = radius * cos_lat[p];
#pragma unroll
    for (int f = 0; f < FT; f += 2) {
      const int level = (c0 + f) >> 1;
      if (level < nout) {
        const double psi_e = o0[f] / east_scale, psi_n = o1[f] / radius;
        const double chi_e = o0[f + 1] / east_scale, chi_n = o1[f + 1] / radius;
        out0[(size_t)level * npts + p] = -psi_n + chi_e;
        out1[(size_t)level * npts + p] = psi_e + chi_n;
      }
    }
  }
}

}  // namespace sht

extern "C" hipError_t sht_launch_sample_pack(const sht::SamplePackArgs* a, hipStream_t s) {
  const long long total = (long long)a->ldc * (a->T + 1) * (a->T + 1);
  if (total <= 0) return hipSuccess;
  long long blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(sht::sample_pack_kernel, dim3((unsigned)blocks), dim3(256), 0, s, *a);
  return hipGetLastError();
}

// Field tile per precision and kind: fp64 accumulators take twice the
// registers of fp32 ones, gradients twice those of values.
int sht_sample_field_tile(int kind, int f32) { return (kind == 0 || f32) ? 8 : 4; }

extern "C" hipError_t sht_launch_sample(const sht::SampleArgs* a, hipStream_t s) {
  if (a->npts <= 0 || a->ldc <= 0) return hipSuccess;
  const int ft = sht_sample_field_tile(a->kind, a->f32);
  if (a->ldc % ft) return hipErrorInvalidValue;
  const dim3 g((unsigned)((a->npts + 255) / 256), (unsigned)(a->ldc / ft)), b(256);
#define SHT_SAMPLE(FT, KIND, F32)                                                             \
  hipLaunchKernelGGL((sht::sample_kernel<FT, KIND, F32>), g, b, 0, s,                          \
                     static_cast<const typename sht::SampleAcc<F32>::C*>(a->coef), a->rec,   \
                     a->diag, a->sin_lat, a->cos_lat, a->lon, a->T, a->npts, a->ldc, a->nout, \
                     a->radius, a->out0, a->out1)
  if (a->f32) {
    if (a->kind == 0) SHT_SAMPLE(8, 0, true);
    else if (a->kind == 1) SHT_SAMPLE(8, 1, true);
    else SHT_SAMPLE(8, 2, true);
  } else {
    if (a->kind == 0) SHT_SAMPLE(8, 0, false);
    else if (a->kind == 1) SHT_SAMPLE(4, 1, false);
    else SHT_SAMPLE(4, 2, false);
  }
#undef SHT_SAMPLE
  return hipGetLastError();
}
