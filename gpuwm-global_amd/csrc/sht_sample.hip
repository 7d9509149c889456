// Spectral fields evaluated at arbitrary points (reference
// spectral/sampling.py:382-515: sample_scalar, sample_gradient, sample_wind).
//
// One lane per two points, one block per (512 points, tile of FT fields, part
// of the order range).  The lane walks its part of the triangle order by order: the normalized associated-Legendre
// recurrence of its point in FP64 (reference legendre.py:86-146, the
// reference's operand order in every expression but the derivative's
// division, see rd below; this file is compiled with -ffp-contract=off and
// the contractions below are explicit fma), and for
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

// The outputs of one point from its two sums (s0: values, or zonal sums
// d/dlambda; s1: meridional sums d/dphi).  KIND 0: the field; 1: east and
// north derivatives per metre; 2: u and v from column pairs (psi, chi):
// u = -north(psi) + east(chi), v = east(psi) + north(chi) (reference
// sampling.py:481-515).
template <int KIND>
__device__ __forceinline__ void sample_store(int col, int p, int npts, int nout, double s0,
                                             double s1, double s0b, double s1b, double radius,
                                             double coslat, double* out0, double* out1) {
  if constexpr (KIND == 0) {
    if (col < nout) out0[(size_t)col * npts + p] = s0;
  } else if constexpr (KIND == 1) {
    if (col < nout) {
      out0[(size_t)col * npts + p] = s0 / (radius * coslat);
      out1[(size_t)col * npts + p] = s1 / radius;
    }
  } else {
    const int level = col >> 1;
    if (level < nout) {
      const double east_scale = radius * coslat;
      const double psi_e = s0 / east_scale, psi_n = s1 / radius;
      const double chi_e = s0b / east_scale, chi_n = s1b / radius;
      out0[(size_t)level * npts + p] = -psi_n + chi_e;
      out1[(size_t)level * npts + p] = psi_e + chi_n;
    }
  }
}

// One block: 256 x PT points (lane l takes points l, l + 256, ...) x FT field
// columns x one part of the order range (orders mcut[r] .. mcut[r+1]-1, cut
// so the parts hold equal shares of the triangle).  With one part the block
// stores the outputs; with several it stores its two raw sums per (column,
// point) into part r of `partial` and sample_reduce_kernel adds the parts in
// order.  The (m, n) walk is one loop over the packed index k.  The
// coefficients arrive through scalar loads; those return out of order, so a
// wave can wait only for all of them, and the next step's loads are the only
// ones in flight while a step computes: each lane takes PT points so that
// every load feeds PT x FT x 2 multiply-adds.
constexpr int kSamplePts = 2;  // PT

template <int FT, int KIND, bool F32>
__global__ void __launch_bounds__(256) sample_kernel(
    const typename SampleAcc<F32>::C* __restrict__ coef, const double* __restrict__ rec,
    const double* __restrict__ diag, const double* __restrict__ sin_lat,
    const double* __restrict__ cos_lat, const double* __restrict__ lon, SampleCuts cuts, int T,
    int npts, int ldc, int nout, double radius, double* __restrict__ out0,
    double* __restrict__ out1, double* __restrict__ partial) {
  using Acc = typename SampleAcc<F32>::T;
  using Cx = typename SampleAcc<F32>::C;
  constexpr bool GRAD = KIND != 0;
  constexpr int PT = kSamplePts;
  const int c0 = blockIdx.y * FT;
  const int r = blockIdx.z;
  const int m0 = cuts.m[r], m1 = cuts.m[r + 1];
  const long long K = (long long)(T + 1) * (T + 2) / 2;
  const double* rec_a = rec;
  const double* rec_b = rec + K;
  const double* rec_c = rec + 2 * K;
  const double* diag_f = diag;
  const double* diag_g = diag + (T + 1);
  const Cx* cf = coef + c0;
  const double pi = 3.14159265358979323846;

  int pt[PT];
  double x[PT], lam[PT], root[PT], denom[PT], dg[PT], prev[PT], prev2[PT];
  // dP = root (n x P - c P_{n-1}) / (x^2 - 1): the division by the point's
  // constant (x^2 - 1) is one multiplication by rd = root / (x^2 - 1), one
  // rounding more than the reference's form and ~10 FP64 instructions less
  // per (n, m) step
  double rd[PT];
#pragma unroll
  for (int q = 0; q < PT; ++q) {
    pt[q] = blockIdx.x * (256 * PT) + q * 256 + threadIdx.x;
    const int pp = pt[q] < npts ? pt[q] : npts - 1;  // idle lanes evaluate a real point
    x[q] = sin_lat[pp];
    lam[q] = lon[pp];
    root[q] = sqrt(fmax(0.0, 1.0 - x[q] * x[q]));
    denom[q] = x[q] * x[q] - 1.0;
    rd[q] = root[q] / denom[q];
    dg[q] = 1.0 / sqrt(4.0 * pi);
    for (int mm = 1; mm <= m0; ++mm) dg[q] = diag_f[mm] * root[q] * dg[q];  // P_{m0 m0}
    prev[q] = prev2[q] = 0.0;
  }

  double o0[PT][FT], o1[PT][FT];
  Acc ar[PT][FT], ai[PT][FT], dr[PT][FT], di[PT][FT];
#pragma unroll
  for (int q = 0; q < PT; ++q)
#pragma unroll
    for (int f = 0; f < FT; ++f) {
      o0[q][f] = o1[q][f] = 0.0;
      ar[q][f] = ai[q][f] = dr[q][f] = di[q][f] = Acc(0);
    }

  // The packed triangle is order-major, so the steps (m, n) of the part are
  // consecutive entries k0 .. k_last: the coefficient and recurrence pointers
  // advance by one entry per step, and each order runs as its diagonal step,
  // its first off-diagonal step and a branch-free loop over n = m+2 .. T.
  // Every step loads the next entry's values once it has used its own (the
  // last step reloads its own, so no load passes k_last).
  const long long k0 = tri_index(T, m0, m0);         // (n, m) = (m0, m0)
  const long long k_last = tri_index(T, T, m1 - 1);  // (T, m1 - 1)
  const Cx* cp = cf + k0 * ldc;
  const double *pa = rec_a + k0, *pb = rec_b + k0, *pc = rec_c + k0;
  long long left = k_last - k0;  // steps after the current one
  Cx cur[FT];
#pragma unroll
  for (int f = 0; f < FT; ++f) cur[f] = cp[f];
  double ca = pa[0], cb = pb[0], cc = pc[0];
  // one step: P (and dP) of every point from `rule`, into the amplitudes
  auto step = [&](auto rule) {
#pragma unroll
    for (int q = 0; q < PT; ++q) {
      double P, dP = 0.0;
      rule(q, P, dP);
#pragma unroll
      for (int f = 0; f < FT; ++f) {
        if constexpr (F32) {
          ar[q][f] = fmaf(cur[f].x, (float)P, ar[q][f]);
          ai[q][f] = fmaf(cur[f].y, (float)P, ai[q][f]);
          if constexpr (GRAD) {
            dr[q][f] = fmaf(cur[f].x, (float)dP, dr[q][f]);
            di[q][f] = fmaf(cur[f].y, (float)dP, di[q][f]);
          }
        } else {
          ar[q][f] = fma(cur[f].x, P, ar[q][f]);
          ai[q][f] = fma(cur[f].y, P, ai[q][f]);
          if constexpr (GRAD) {
            dr[q][f] = fma(cur[f].x, dP, dr[q][f]);
            di[q][f] = fma(cur[f].y, dP, di[q][f]);
          }
        }
      }
      prev2[q] = prev[q];
      prev[q] = P;
    }
    // the next entry's values go into the same registers once this step has
    // used them: scalar loads return out of order, so a wave waits for all of
    // its outstanding ones at once and a second register set would buy no
    // overlap, only scalar registers (and their spills)
    const long long adv = left > 0 ? 1 : 0;
    cp += adv * ldc;
    pa += adv;
    pb += adv;
    pc += adv;
    --left;
#pragma unroll
    for (int f = 0; f < FT; ++f) cur[f] = cp[f];
    ca = pa[0];
    cb = pb[0];
    cc = pc[0];
  };
  for (int m = m0; m < m1; ++m) {
    // n = m
    step([&](int q, double& P, double& dP) {
      P = dg[q];
      if constexpr (GRAD) dP = (m >= 1) ? rd[q] * ((double)m * x[q] * dg[q]) : 0.0;
    });
    if (m < T) {
      // n = m + 1
      const double g = diag_g[m];
      const int n = m + 1;
      step([&](int q, double& P, double& dP) {
        P = g * x[q] * dg[q];
        if constexpr (GRAD) dP = rd[q] * ((double)n * x[q] * P - cc * prev[q]);
      });
      for (int n2 = m + 2; n2 <= T; ++n2) {
        step([&](int q, double& P, double& dP) {
          P = ca * (x[q] * prev[q] - cb * prev2[q]);
          if constexpr (GRAD) dP = rd[q] * ((double)n2 * x[q] * P - cc * prev[q]);
        });
      }
    }
    // the order is complete: turn its amplitudes by e^{i m lambda}
    const double mult = m == 0 ? 1.0 : 2.0;
#pragma unroll
    for (int q = 0; q < PT; ++q) {
      double sm, cm;
      sincos((double)m * lam[q], &sm, &cm);
#pragma unroll
      for (int f = 0; f < FT; ++f) {
        const double re = (double)ar[q][f], im = (double)ai[q][f];
        if constexpr (KIND == 0) {
          o0[q][f] += mult * (re * cm - im * sm);
        } else {
          // Re(i m A e^{i m lambda}) and Re(dA e^{i m lambda})
          o0[q][f] += mult * (-(double)m * (re * sm + im * cm));
          o1[q][f] += mult * ((double)dr[q][f] * cm - (double)di[q][f] * sm);
        }
        ar[q][f] = ai[q][f] = dr[q][f] = di[q][f] = Acc(0);
      }
    }
    if (m + 1 <= T) {
      const double fm = diag_f[m + 1];
#pragma unroll
      for (int q = 0; q < PT; ++q) dg[q] = fm * root[q] * dg[q];
    }
  }
#pragma unroll
  for (int q = 0; q < PT; ++q) {
    const int p = pt[q];
    if (p >= npts) continue;
    if (gridDim.z > 1) {
      // raw sums of this part: partial[r][c][p], then [R + r] for the second sum
      const size_t plane = (size_t)ldc * npts;
#pragma unroll
      for (int f = 0; f < FT; ++f) {
        partial[(size_t)r * plane + (size_t)(c0 + f) * npts + p] = o0[q][f];
        if constexpr (GRAD)
          partial[(size_t)(gridDim.z + r) * plane + (size_t)(c0 + f) * npts + p] = o1[q][f];
      }
      continue;
    }
    const double coslat = KIND == 0 ? 0.0 : cos_lat[p];
    if constexpr (KIND == 2) {
      static_assert(FT % 2 == 0, "wind tiles hold whole (psi, chi) pairs");
#pragma unroll
      for (int f = 0; f < FT; f += 2)
        sample_store<KIND>(c0 + f, p, npts, nout, o0[q][f], o1[q][f], o0[q][f + 1],
                           o1[q][f + 1], radius, coslat, out0, out1);
    } else {
#pragma unroll
      for (int f = 0; f < FT; ++f)
        sample_store<KIND>(c0 + f, p, npts, nout, o0[q][f], o1[q][f], 0.0, 0.0, radius, coslat,
                           out0, out1);
    }
  }
}

// Adds the R parts' sums of each (column, point) in part order and stores
// the outputs; one thread per (column or column pair, point).
template <int KIND>
__global__ void sample_reduce_kernel(const double* __restrict__ partial, int nparts, int ldc,
                                     int npts, int nout, double radius,
                                     const double* __restrict__ cos_lat, double* out0,
                                     double* out1) {
  constexpr int STRIDE = KIND == 2 ? 2 : 1;
  const long long total = (long long)(ldc / STRIDE) * npts;
  const size_t plane = (size_t)ldc * npts;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int p = (int)(e % npts);
    const int col = (int)(e / npts) * STRIDE;
    if ((KIND == 2 ? col >> 1 : col) >= nout) continue;
    double s[4] = {0.0, 0.0, 0.0, 0.0};  // s0, s1 of col; s0, s1 of col + 1 (winds)
    for (int r = 0; r < nparts; ++r) {
      const size_t o = (size_t)r * plane + (size_t)col * npts + p;
      s[0] += partial[o];
      if (KIND != 0) s[1] += partial[(size_t)nparts * plane + o];
      if (KIND == 2) {
        s[2] += partial[o + npts];
        s[3] += partial[(size_t)nparts * plane + o + npts];
      }
    }
    sample_store<KIND>(col, p, npts, nout, s[0], s[1], s[2], s[3], radius,
                       KIND == 0 ? 0.0 : cos_lat[p], out0, out1);
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
// Fields per block: eight in every mode (the FP64 gradient and wind kernels
// hold 255 VGPRs at eight, two waves per SIMD, and are still 7% faster than
// at four fields with three)
int sht_sample_field_tile(int kind, int f32) {
  (void)kind;
  (void)f32;
  return 8;
}

extern "C" hipError_t sht_launch_sample(const sht::SampleArgs* a, hipStream_t s) {
  if (a->npts <= 0 || a->ldc <= 0) return hipSuccess;
  const int ft = sht_sample_field_tile(a->kind, a->f32);
  const int R = a->cuts.parts;
  if (a->ldc % ft || R < 1 || R > sht::kSampleMaxParts || (R > 1 && !a->partial))
    return hipErrorInvalidValue;
  constexpr int kPointsPerBlock = 256 * sht::kSamplePts;
  const dim3 g((unsigned)((a->npts + kPointsPerBlock - 1) / kPointsPerBlock),
               (unsigned)(a->ldc / ft), (unsigned)R),
      b(256);
#define SHT_SAMPLE(FT, KIND, F32)                                                             \
  hipLaunchKernelGGL((sht::sample_kernel<FT, KIND, F32>), g, b, 0, s,                          \
                     static_cast<const typename sht::SampleAcc<F32>::C*>(a->coef), a->rec,   \
                     a->diag, a->sin_lat, a->cos_lat, a->lon, a->cuts, a->T, a->npts, a->ldc, \
                     a->nout, a->radius, a->out0, a->out1, a->partial)
  if (a->f32) {
    if (a->kind == 0) SHT_SAMPLE(8, 0, true);
    else if (a->kind == 1) SHT_SAMPLE(8, 1, true);
    else SHT_SAMPLE(8, 2, true);
  } else {
    if (a->kind == 0) SHT_SAMPLE(8, 0, false);
    else if (a->kind == 1) SHT_SAMPLE(8, 1, false);
    else SHT_SAMPLE(8, 2, false);
  }
#undef SHT_SAMPLE
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || R == 1) return e;
  const int stride = a->kind == 2 ? 2 : 1;
  long long blocks = ((long long)(a->ldc / stride) * a->npts + 255) / 256;
  if (blocks > 65536) blocks = 65536;
#define SHT_REDUCE(KIND)                                                                       \
  hipLaunchKernelGGL(sht::sample_reduce_kernel<KIND>, dim3((unsigned)blocks), dim3(256), 0, s,  \
                     a->partial, R, a->ldc, a->npts, a->nout, a->radius, a->cos_lat, a->out0, \
                     a->out1)
  if (a->kind == 0) SHT_REDUCE(0);
  else if (a->kind == 1) SHT_REDUCE(1);
  else SHT_REDUCE(2);
#undef SHT_REDUCE
  return hipGetLastError();
}
