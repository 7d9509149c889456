// Gaussian nodes and associated-Legendre tables, generated on the device in FP64.
//
// Nodes: Newton iteration on P_nlat(x) (the classic Gauss-Legendre root
// finder), one thread per positive root; weights w = 2 / ((1-x^2) P'_nlat(x)^2).
// The reference obtains the same nodes from numpy's leggauss
// (reference spectral/grid.py:112); they agree to a few ulp.
//
// Legendre functions: the normalized recurrence of reference
// spectral/legendre.py:86-113 and the meridional derivative of
// legendre.py:120-146, with the reference's operand order in every
// expression.  This file is compiled with -ffp-contract=off so no
// multiply-add is fused: given the same nodes the FP64 values are the
// reference's bit for bit.
//
// One thread owns one (order m, folded row i) pair and walks n = m..T, so the
// O(T) diagonal product is recomputed per thread; the whole T767 table set
// costs ~1e8 FP64 operations, negligible beside one transform.
#include "sht_kernels.h"

#include <math.h>

namespace sht {

// P_nlat(x) and P_{nlat-1}(x) by the three-term recurrence.
__device__ static inline void legendre_pair(int n, double x, double* pn, double* pn1) {
  double p0 = 1.0, p1 = x;
  if (n == 0) { *pn = 1.0; *pn1 = 0.0; return; }
  for (int k = 2; k <= n; ++k) {
    double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  *pn = p1;
  *pn1 = p0;
}

// mu[] ascending (south to north) and weights, nlat >= 1.
__global__ void gauss_newton_kernel(int nlat, double* __restrict__ mu,
                                    double* __restrict__ w) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;  // k-th largest root
  int half = nlat / 2;
  if (k >= (nlat + 1) / 2) return;
  if ((nlat & 1) && k == half) {  // equator root of an odd-order polynomial
    double pn, pn1;
    legendre_pair(nlat, 0.0, &pn, &pn1);
    double dp = nlat * (0.0 * pn - pn1) / (0.0 - 1.0);
    mu[half] = 0.0;
    w[half] = 2.0 / (dp * dp);
    return;
  }
  const double pi = 3.14159265358979323846;
  double x = cos(pi * (k + 0.75) / (nlat + 0.5));
  double dp = 1.0;
  for (int it = 0; it < 100; ++it) {
    double pn, pn1;
    legendre_pair(nlat, x, &pn, &pn1);
    dp = nlat * (x * pn - pn1) / (x * x - 1.0);
    double dx = pn / dp;
    x -= dx;
    if (fabs(dx) <= 4e-16 * fabs(x)) break;
  }
  {  // one polish step and the derivative at the final node
    double pn, pn1;
    legendre_pair(nlat, x, &pn, &pn1);
    dp = nlat * (x * pn - pn1) / (x * x - 1.0);
    x -= pn / dp;
    legendre_pair(nlat, x, &pn, &pn1);
    dp = nlat * (x * pn - pn1) / (x * x - 1.0);
  }
  double wk = 2.0 / ((1.0 - x * x) * dp * dp);
  mu[nlat - 1 - k] = x;
  mu[k] = -x;
  w[nlat - 1 - k] = wk;
  w[k] = wk;
}

// Round-to-nearest-even float -> bf16 bits.
__device__ static inline unsigned short bf16_bits(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ static inline float bf16_value(unsigned short b) {
  return __uint_as_float(((unsigned int)b) << 16);
}



// f32 and bf16 tables use the k-chunk tiled layouts; the FP64 copy (tests)
// keeps the plain synthesis layout, row i then degree.
__device__ static inline void store_value(const TableOut& out, int kind, long long plain_idx,
                                          long long syn_idx, long long ana_idx, double v) {
  if (out.f32_syn[kind]) out.f32_syn[kind][syn_idx] = (float)v;
  if (out.f32_ana[kind]) out.f32_ana[kind][ana_idx] = (float)v;
  if (out.f64_syn[kind]) out.f64_syn[kind][plain_idx] = v;
  if (out.bf_syn[kind] || out.bf_ana[kind]) {
    unsigned short b0 = bf16_bits((float)v);
    double r1 = v - (double)bf16_value(b0);
    unsigned short b1 = bf16_bits((float)r1);
    double r2 = r1 - (double)bf16_value(b1);
    unsigned short b2 = bf16_bits((float)r2);
    if (out.bf_syn[kind]) {
      out.bf_syn[kind][syn_idx] = b0;
      out.bf_syn[kind][out.elems + syn_idx] = b1;
      out.bf_syn[kind][2 * out.elems + syn_idx] = b2;
    }
    if (out.bf_ana[kind]) {
      out.bf_ana[kind][ana_idx] = b0;
      out.bf_ana[kind][out.elems + ana_idx] = b1;
      out.bf_ana[kind][2 * out.elems + ana_idx] = b2;
    }
  }
}

// koff[2m+p] = first packed k of block (m,p); the table block starts at
// koff * JHP in either layout.
__global__ void legendre_tables_kernel(int T, int nlat, int JH, int JHP,
                                       const double* __restrict__ mu,
                                       const int* __restrict__ koff, TableOut out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int m = blockIdx.y;
  if (i >= JH || m > T) return;
  const double x = mu[nlat - 1 - i];  // north node, x >= 0
  const double root = sqrt(fmax(0.0, 1.0 - x * x));
  const double denom = x * x - 1.0;
  const double cosl = root;
  const double pi = 3.14159265358979323846;
  double diag = 1.0 / sqrt(4.0 * pi);
  for (int mm = 1; mm <= m; ++mm) diag = -sqrt((2 * mm + 1) / (2.0 * mm)) * root * diag;

  double p2 = 0.0, p1 = 0.0;  // P_{n-2}, P_{n-1}
  // degrees m..T, and T+1 of the basis alone: the vector operators'
  // recurrence form contracts the basis one degree past the truncation
  for (int n = m; n <= T + 1; ++n) {
    double p;
    if (n == m) {
      p = diag;
    } else if (n == m + 1) {
      p = sqrt((double)(2 * m + 3)) * x * diag;
    } else {
      double a = sqrt((4.0 * n * n - 1.0) / (double)(n * n - m * m));
      double b = sqrt(((n - 1.0) * (n - 1.0) - (double)(m * m)) /
                      (4.0 * (n - 1.0) * (n - 1.0) - 1.0));
      p = a * (x * p1 - b * p2);
    }
    double d;
    if (n == m) {
      d = (m >= 1) ? cosl * (m * x * p) / denom : 0.0;
    } else {
      double c = sqrt((double)(n * n - m * m) * (2.0 * n + 1.0) / (2.0 * n - 1.0));
      double num = n * x * p - c * p1;
      d = cosl * num / denom;
    }
    double pm = (m == 0) ? 0.0 : (p * (double)m) / cosl;
    const int par = (n - m) & 1;
    const int nidx = (n - m) >> 1;
    const long long base = (long long)koff[2 * m + par] * JHP;
    const int kp = koff[2 * m + par + 1] - koff[2 * m + par];
    const long long plain_idx = base + (long long)i * kp + nidx;
    const long long syn_idx = base + syn_tiled(JHP, i, nidx);
    const long long ana_idx = base + ana_tiled(kp, i, nidx);
    if (m < out.low_m0) {  // FP64 basis of the low orders, degrees m..T+1
      if (out.f64_low_syn[0]) out.f64_low_syn[0][((long long)m * (T + 2) + (n - m)) * JHP + i] = p;
      if (out.f64_low_ana[0]) out.f64_low_ana[0][((long long)m * JHP + i) * (T + 2) + (n - m)] = p;
    }
    if (n > T) {  // the extra degree: P only (the other kinds keep their zeros)
      if (nidx < kp) store_value(out, 0, plain_idx, syn_idx, ana_idx, p);
      break;
    }
    store_value(out, 0, plain_idx, syn_idx, ana_idx, p);
    store_value(out, 1, plain_idx, syn_idx, ana_idx, d);
    store_value(out, 2, plain_idx, syn_idx, ana_idx, pm);
    if (m < out.low_m0) {
      const int r = n - m;
      const long long ls = ((long long)m * (T + 1) + r) * JHP + i;
      const long long la = ((long long)m * JHP + i) * (T + 1) + r;
      const double vk[3] = {p, d, pm};
#pragma unroll
      for (int k = 1; k < 3; ++k) {  // (kind 0: the basis above, T+2 rows per order)
        if (out.f64_low_syn[k]) out.f64_low_syn[k][ls] = vk[k];
        if (out.f64_low_ana[k]) out.f64_low_ana[k][la] = vk[k];
      }
    }
    p2 = p1;
    p1 = p;
  }
}

}  // namespace sht

extern "C" hipError_t sht_launch_gauss_newton(int nlat, double* mu, double* w, hipStream_t s) {
  int roots = (nlat + 1) / 2;
  hipLaunchKernelGGL(sht::gauss_newton_kernel, dim3((roots + 63) / 64), dim3(64), 0, s,
                     nlat, mu, w);
  return hipGetLastError();
}

extern "C" hipError_t sht_launch_legendre_tables(int T, int nlat, int JH, int JHP,
                                                 const double* mu, const int* koff,
                                                 const sht::TableOut* out, hipStream_t s) {
  dim3 grid((JH + 63) / 64, T + 1);
  hipLaunchKernelGGL(sht::legendre_tables_kernel, grid, dim3(64), 0, s, T, nlat, JH, JHP, mu,
                     koff, *out);
  return hipGetLastError();
}
