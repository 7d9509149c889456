// The vector operators' FP64 low orders (uv_from_vordiv, vordiv_from_uv).
//
// A wind field's large scales sit in its lowest zonal orders.  Vorticity and
// divergence at high degree are small differences of those large values
// weighted by dP/dphi (which grows like n), so fp32 rounding anywhere on the
// low orders (FFT, waist storage, table values, accumulation) is amplified
// by ~T/m relative to the field's largest coefficient.  For the orders
// m < m0 the engine therefore runs the whole Legendre stage in FP64: these
// kernels contract the FP64 tables of those orders with the FP64 low waist
// (written by the FP64 fold, read by the FP64 unfold, sht_fft.hip), while
// the orders m >= m0 go through the grouped GEMM as before.
//
// Low waist of a vector job: [level][f][m][p][i] double2 (complex), f = 0
// for u and 1 for v, p the parity, i the folded row.  Folded row i pairs the
// north row nlat-1-i with the south row i; E (p = 0) and O (p = 1) are the
// even and odd parts, north = E + O, south = E - O.  A table of parity
// (n - m) mod 2 = q contributes to E when q = 0; dP/dphi has the opposite
// parity of P and P m / cos.
//
// Synthesis (reference vector.py:36-52, transform.py:1748-1768): the winds
// times a cos(phi) as one contraction of the basis over degrees m..T+1 (the
// recurrence form, below; the unfold applies 1/(a cos)).  Analysis
// (vector.py:113-134), with the FP64 fold's weights 2 pi w / (a nlon)
// already in E and O:
//   zeta = i sum Pm V_q + sum dP U_(1-q),   D = i sum Pm U_q - sum dP V_(1-q).
#include <algorithm>

#include "sht_kernels.h"

namespace sht {

// Both directions are small FP64 GEMMs per order m, run on FP64 MFMA
// (v_mfma_f64_16x16x4_f64: A 16 x 4 with lane l holding A[l % 16][l / 16],
// B 4 x 16 with lane l holding B[l / 16][l % 16], D 16 x 16 with lane l,
// register r holding D[l / 16 + 4 r][l % 16] -- rows interleaved, unlike the
// fp32 16x16x4 MFMA's 4 (l / 16) + r; checked on gfx950).  A block owns one
// order m, one output parity, 8 levels (32 data columns: two 16-column MFMA
// tiles) and up to 4 row tiles of 16, two wavefronts per row tile: each takes
// one half of every segment's k range, and the halves' sums meet in LDS at
// the end (the grids are a few hundred blocks, so the k loop's length, not
// the MFMA rate, sets the time).  The B operand of one segment (table kind)
// at a time is staged in LDS in FP64; each A value a wavefront loads from its
// table feeds both column tiles.
typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kLowCols = 32;    // data columns per block: 8 levels x 4 components
constexpr int kLowLevels = 8;
constexpr int kLowWaves = 8;    // wavefronts per block
constexpr int kLowTiles = kLowWaves / 2;  // row tiles per block (two wavefronts each)
constexpr int kLowBatch = 8;    // k-steps of A loaded ahead (32 k)
constexpr int kLowKPad = 4 * kLowBatch;  // k extents padded to whole batches
// LDS for the two halves' sums: per row tile, 64 lanes x two d4 accumulators
constexpr size_t kLowReduceBytes = (size_t)kLowTiles * 2 * 64 * 4 * sizeof(double);

// Where the second half of a k range of kpad (a multiple of kLowKPad) starts.
__device__ __forceinline__ int low_split(int kpad) {
  return (kpad / kLowKPad + 1) / 2 * kLowKPad;
}

__device__ __forceinline__ size_t low_index(int JHP, int m0, int lev, int f, int m, int p, int i) {
  return ((((size_t)lev * 2 + f) * m0 + m) * 2 + p) * JHP + i;
}

// acc0 += A B[:, 0:16], acc1 += A B[:, 16:32] over kpad k (multiple of
// kLowKPad); a(k) gives this lane's A value for k-step row k (k = j0 + lane/16
// + 4 u), any finite value where the B row is zero.  A values come from
// global memory 8 k-steps at a time, the next batch in flight while the
// current one feeds the MFMAs.
template <class LoadA>
__device__ __forceinline__ void low_segment(d4& acc0, d4& acc1, const double* b, int kpad,
                                            int lane, LoadA&& a) {
  const int kq = lane >> 4, col = lane & 15;
  double x[kLowBatch], y[kLowBatch];  // two batches in flight, alternating roles
  auto load = [&](double (&v)[kLowBatch], int j0) {
#pragma unroll
    for (int u = 0; u < kLowBatch; ++u) v[u] = a(j0 + 4 * u + kq);
  };
  auto mma = [&](const double (&v)[kLowBatch], int j0) {
    // the batch's B values first (one LDS wait), then its MFMAs
    double b0[kLowBatch], b1[kLowBatch];
#pragma unroll
    for (int u = 0; u < kLowBatch; ++u) {
      const double* row = b + (size_t)(j0 + 4 * u + kq) * kLowCols;
      b0[u] = row[col];
      b1[u] = row[16 + col];
    }
#pragma unroll
    for (int u = 0; u < kLowBatch; ++u) {
      acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(v[u], b0[u], acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(v[u], b1[u], acc1, 0, 0, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x100, 2 * kLowBatch, 0);  // the LDS reads first
    __builtin_amdgcn_sched_group_barrier(0x008, 2 * kLowBatch, 0);  // then the MFMAs
  };
  load(x, 0);
  for (int j0 = 0; j0 < kpad; j0 += 2 * kLowKPad) {
    load(y, j0 + kLowKPad);      // past the end: clamped rows, never multiplied
    mma(x, j0);
    load(x, j0 + 2 * kLowKPad);
    if (j0 + kLowKPad < kpad) mma(y, j0 + kLowKPad);
  }
}

// The second wavefront of each row tile (kh = 1) hands its sums to the first
// through LDS (after every wavefront is done with the staged B); returns
// whether this wavefront holds the tile's totals and stores them.
__device__ __forceinline__ bool low_reduce(d4& acc0, d4& acc1, double* lds, int w, int lane,
                                           bool active) {
  const int kh = w & 1;
  d4* red = reinterpret_cast<d4*>(lds) + (size_t)(w >> 1) * 2 * 64;
  __syncthreads();
  if (kh) {
    red[lane] = acc0;
    red[64 + lane] = acc1;
  }
  __syncthreads();
  if (kh || !active) return false;
  acc0 += red[lane];
  acc1 += red[64 + lane];
  return true;
}

// Synthesis in the recurrence form (as the GEMM's orders above m0; sht_pack.hip
// states the identity): output parity p, rows i (folded latitude), columns
// (level, component c), c = 0, 1 the real and imaginary part of a cos(phi) u,
// c = 2, 3 of a cos(phi) v:
//   out[i][col] = sum_{r of parity p} P[r][i] W[r][col],  W = (U, V) over
//   degrees k = m + r <= T + 1,
//   U_k = (k-1) e_k psi_{k-1} - (k+2) e_{k+1} psi_{k+1} + i m chi_k,
//   V_k = i m psi_k - (k-1) e_k chi_{k-1} + (k+2) e_{k+1} chi_{k+1},
// formed in FP64 from psi, chi = inverse Laplacian (vorticity, divergence) as
// the block stages them.  In FP64 the order-0 cancellation near the poles
// (terms O(1), sum O(cos^2)) costs ~1/cos x 1e-16, so every order m < m0
// takes this form: one table segment instead of the P m / cos and dP/dphi
// pair.  Grid: x = m0 x row-tile groups, y = level groups of 8, z = parity.
__global__ void __launch_bounds__(64 * kLowWaves) low_synthesis_kernel(LowSynArgs a) {
  extern __shared__ double lds_low[];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int groups = (a.JH + 16 * kLowTiles - 1) / (16 * kLowTiles);
  const int m = blockIdx.x / groups;
  const int it = (blockIdx.x % groups) * kLowTiles + (w >> 1);  // this wavefront's row tile
  const int kh = w & 1;                                          // and its half of k
  const bool active = it * 16 < a.JH;                            // wave-uniform
  const int lev0 = blockIdx.y * kLowLevels;
  const int p = blockIdx.z;
  const int T1 = a.T + 1;
  const int kn = (a.T + 3 - m - p) / 2;  // degrees k = m + p + 2 j <= T + 1
  const int kpad = (kn + kLowKPad - 1) / kLowKPad * kLowKPad;
  const int i = it * 16 + (lane & 15);
  const int ic = min(i, a.JHP - 1);
  d4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
  // one thread per (level, k row): psi, chi of degrees k - 1, k, k + 1 give
  // U_k, V_k of that level; all loads of a thread are issued before the LDS
  // writes
  constexpr int kPer = 2;
  const double a2 = a.radius * a.radius;
  const double dm = (double)m;
  for (int e0 = threadIdx.x; e0 < kpad * kLowLevels; e0 += kPer * blockDim.x) {
    float2 zv[kPer][3], dv[kPer][3];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = e0 + u * blockDim.x;
      const int l = e % kLowLevels, j = e / kLowLevels;
      const int lev = min(lev0 + l, a.nlev - 1);
      const int k = m + p + 2 * j;
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        const int n = min(max(k - 1 + s, m), a.T);  // clamped: never used outside m..T
        const size_t o = ((size_t)lev * T1 + n) * T1 + m;
        zv[u][s] = a.vort[o];
        dv[u][s] = a.div[o];
      }
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = e0 + u * blockDim.x;
      const int l = e % kLowLevels, j = e / kLowLevels;
      if (j >= kpad) continue;
      const int k = m + p + 2 * j;
      const bool ok = j < kn && lev0 + l < a.nlev;
      double ps[3][2], ch[3][2];  // psi, chi of degrees k - 1, k, k + 1 (re, im)
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        const int n = k - 1 + s;
        const bool in = ok && n >= m && n <= a.T && n >= 1;
        const double f = in ? -a2 / ((double)n * ((double)n + 1.0)) : 0.0;
        ps[s][0] = f * (double)zv[u][s].x;
        ps[s][1] = m == 0 ? 0.0 : f * (double)zv[u][s].y;
        ch[s][0] = f * (double)dv[u][s].x;
        ch[s][1] = m == 0 ? 0.0 : f * (double)dv[u][s].y;
      }
      const double dk = (double)k;
      const double e0k = k > m ? sqrt((dk * dk - dm * dm) / (4.0 * dk * dk - 1.0)) : 0.0;
      const double e1k = sqrt(((dk + 1.0) * (dk + 1.0) - dm * dm) /
                              (4.0 * (dk + 1.0) * (dk + 1.0) - 1.0));
      const double ca = (dk - 1.0) * e0k, cb = (dk + 2.0) * e1k;
      double* dst = lds_low + (size_t)j * kLowCols + 4 * l;
      dst[0] = ca * ps[0][0] - cb * ps[2][0] - dm * ch[1][1];  // U
      dst[1] = ca * ps[0][1] - cb * ps[2][1] + dm * ch[1][0];
      dst[2] = -dm * ps[1][1] - ca * ch[0][0] + cb * ch[2][0];  // V
      dst[3] = dm * ps[1][0] - ca * ch[0][1] + cb * ch[2][1];
    }
  }
  __syncthreads();
  if (active && kn > 0) {
    const int k0 = kh ? low_split(kpad) : 0, k1 = kh ? kpad : low_split(kpad);
    const double* tab = a.p + (size_t)m * (T1 + 1) * a.JHP + ic;
    // rows past the order's degrees meet zero B rows, so any finite table
    // value will do there: clamped addresses, no select (a select right
    // after a load makes the compiler wait for it before issuing the next)
    low_segment(acc0, acc1, lds_low + (size_t)k0 * kLowCols, k1 - k0, lane, [&](int j) {
      return tab[(size_t)min(2 * (j + k0) + p, T1) * a.JHP];
    });
  }
  if (!low_reduce(acc0, acc1, lds_low, w, lane, active)) return;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int col = 16 * h + (lane & 15);
    const int lev = lev0 + col / 4, c = col % 4;
    if (lev >= a.nlev) continue;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int io = it * 16 + (lane >> 4) + 4 * q;
      if (io < a.JH) {
        double* dst = reinterpret_cast<double*>(
            a.low + low_index(a.JHP, a.m0, lev, c >> 1, m, p, io));
        dst[c & 1] = h ? acc1[q] : acc0[q];
      }
    }
  }
}

// Analysis in the recurrence form: output parity q, rows r = 2 j + q (degrees
// k = m + r <= T + 1), columns (level, component c), c = 0, 1 the real and
// imaginary part of A, c = 2, 3 of B:
//   A_k = sum_i P[r][i] U_q[i],   B_k = sum_i P[r][i] V_q[i]
// with U, V the FP64 fold's (E, O) of u and v, weighted by
// 2 pi w / (a nlon cos(phi)).  A and B go to the job's FP64 buffer, and the
// combination below forms vorticity and divergence from neighbouring degrees
// in FP64 (as unpack_vector_rec_kernel does for the other orders): one table
// segment instead of the P m / cos and dP/dphi pair.  Grid: x = m0 x row-tile
// groups, y = level groups of 8, z = parity.
__global__ void __launch_bounds__(64 * kLowWaves) low_analysis_kernel(LowAnaArgs a) {
  extern __shared__ double lds_low[];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int T2 = a.T + 2;
  const int rows_max = (a.T + 3) / 2;                     // rows of parity 0 at m = 0
  const int groups = (rows_max + 16 * kLowTiles - 1) / (16 * kLowTiles);
  const int m = blockIdx.x / groups;
  const int it = (blockIdx.x % groups) * kLowTiles + (w >> 1);
  const int kh = w & 1;
  const int lev0 = blockIdx.y * kLowLevels;
  const int q = blockIdx.z;
  const int KQ = (a.T + 3 - m - q) / 2;                   // rows of parity q
  const bool active = it * 16 < KQ;                       // wave-uniform
  const int IP = (a.JH + kLowKPad - 1) / kLowKPad * kLowKPad;
  const int j = it * 16 + (lane & 15);
  const int r = 2 * min(j, max(KQ - 1, 0)) + q;
  d4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
  {
    // one thread per (level, folded row): U and V of parity q give the four
    // columns of that level; loads first
    constexpr int kPer = 4;
    for (int e0 = threadIdx.x; e0 < IP * kLowLevels; e0 += kPer * blockDim.x) {
      double2 uu[kPer], vv[kPer];
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int e = e0 + u * blockDim.x;
        const int i = min(e % IP, a.JH - 1), l = e / IP;
        const int lev = min(lev0 + l, a.nlev - 1);
        uu[u] = a.low[low_index(a.JHP, a.m0, lev, 0, m, q, i)];
        vv[u] = a.low[low_index(a.JHP, a.m0, lev, 1, m, q, i)];
      }
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int e = e0 + u * blockDim.x;
        const int i = e % IP, l = e / IP;
        if (l >= kLowLevels) continue;
        const bool ok = i < a.JH && lev0 + l < a.nlev;
        const double2 U = ok ? uu[u] : make_double2(0.0, 0.0);
        const double2 V = ok ? vv[u] : make_double2(0.0, 0.0);
        double* dst = lds_low + (size_t)i * kLowCols + 4 * l;
        dst[0] = U.x;
        dst[1] = U.y;
        dst[2] = V.x;
        dst[3] = V.y;
      }
    }
  }
  __syncthreads();
  if (active) {
    const int k0 = kh ? low_split(IP) : 0, k1 = kh ? IP : low_split(IP);
    const double* tab = a.p + (size_t)m * a.JHP * T2 + r;
    low_segment(acc0, acc1, lds_low + (size_t)k0 * kLowCols, k1 - k0, lane,
                [&](int i) {  // past JH: zero B rows
                  return tab[(size_t)min(i + k0, a.JH - 1) * T2];
                });
  }
  if (!low_reduce(acc0, acc1, lds_low, w, lane, active)) return;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int col = 16 * h + (lane & 15);
    const int lev = lev0 + col / 4, c = col % 4;
    if (lev >= a.nlev) continue;
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) {
      const int jo = it * 16 + (lane >> 4) + 4 * qq;
      if (jo < KQ) {
        double* dst = reinterpret_cast<double*>(
            a.ab + (((size_t)lev * 2 + (c >> 1)) * T2 + m + 2 * jo + q) * a.m0 + m);
        dst[c & 1] = h ? acc1[qq] : acc0[qq];
      }
    }
  }
}

// Vorticity and divergence of the orders m < m0 from the FP64 A, B of the
// analysis above (degrees n - 1, n, n + 1; the unpack's recurrence sums for
// the other orders, here in FP64 throughout):
//   zeta_n = i m B_n - n e_{n+1} A_{n+1} + (n+1) e_n A_{n-1}
//   D_n    = i m A_n + n e_{n+1} B_{n+1} - (n+1) e_n B_{n-1}
// One thread per (level, n, m < m0), m fastest (consecutive A, B entries).
__global__ void __launch_bounds__(256) low_combine_kernel(LowAnaArgs a) {
  const int T1 = a.T + 1, T2 = a.T + 2, M0 = min(a.m0, T1);
  const long long total = (long long)a.nlev * T1 * M0;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int m = (int)(e % M0);
  const int n = (int)((e / M0) % T1);
  const int lev = (int)(e / ((long long)M0 * T1));
  if (n < m) return;  // above the triangle: the unpack writes those zeros
  const double2* A = a.ab + (size_t)lev * 2 * T2 * a.m0 + m;
  const double2* B = A + (size_t)T2 * a.m0;
  const double2 an = A[(size_t)n * a.m0], bn = B[(size_t)n * a.m0];
  const double2 ap = A[(size_t)(n + 1) * a.m0], bp = B[(size_t)(n + 1) * a.m0];
  const double2 am = n > m ? A[(size_t)(n - 1) * a.m0] : make_double2(0.0, 0.0);
  const double2 bm = n > m ? B[(size_t)(n - 1) * a.m0] : make_double2(0.0, 0.0);
  const double dn = (double)n, dm = (double)m;
  auto eps = [](double nn, double mm) {
    return nn > mm ? sqrt((nn * nn - mm * mm) / (4.0 * nn * nn - 1.0)) : 0.0;
  };
  const double ca = (dn + 1.0) * eps(dn, dm), cb = dn * eps(dn + 1.0, dm);
  float2 z = make_float2((float)(-dm * bn.y - cb * ap.x + ca * am.x),
                         (float)(dm * bn.x - cb * ap.y + ca * am.y));
  float2 d = make_float2((float)(-dm * an.y + cb * bp.x - ca * bm.x),
                         (float)(dm * an.x + cb * bp.y - ca * bm.y));
  if (m == 0) {
    z.y = 0.f;
    d.y = 0.f;
  }
  const size_t o = ((size_t)lev * T1 + n) * T1 + m;
  a.vort[o] = z;
  a.div[o] = d;
}

}  // namespace sht

extern "C" hipError_t sht_launch_low_synthesis(const sht::LowSynArgs* a, hipStream_t s) {
  if (a->nlev <= 0 || a->m0 <= 0) return hipSuccess;
  const int m0 = std::min(a->m0, a->T + 1);
  const int groups = (a->JH + 16 * sht::kLowTiles - 1) / (16 * sht::kLowTiles);
  const int KP = ((a->T + 3) / 2 + sht::kLowKPad - 1) / sht::kLowKPad * sht::kLowKPad;
  const size_t lds = std::max((size_t)KP * sht::kLowCols * sizeof(double),  // the B operand
                              sht::kLowReduceBytes);
  const dim3 g(m0 * groups, (a->nlev + sht::kLowLevels - 1) / sht::kLowLevels, 2);
  hipLaunchKernelGGL(sht::low_synthesis_kernel, g, dim3(64 * sht::kLowWaves), lds, s, *a);
  return hipGetLastError();
}

extern "C" hipError_t sht_launch_low_analysis(const sht::LowAnaArgs* a, hipStream_t s) {
  if (a->nlev <= 0 || a->m0 <= 0) return hipSuccess;
  const int m0 = std::min(a->m0, a->T + 1);
  const int rows_max = (a->T + 3) / 2;  // degrees up to T+1
  const int groups = (rows_max + 16 * sht::kLowTiles - 1) / (16 * sht::kLowTiles);
  const int IP = (a->JH + sht::kLowKPad - 1) / sht::kLowKPad * sht::kLowKPad;
  const size_t lds = std::max((size_t)IP * sht::kLowCols * sizeof(double), sht::kLowReduceBytes);
  const dim3 g(m0 * groups, (a->nlev + sht::kLowLevels - 1) / sht::kLowLevels, 2);
  hipLaunchKernelGGL(sht::low_analysis_kernel, g, dim3(64 * sht::kLowWaves), lds, s, *a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const long long total = (long long)a->nlev * (a->T + 1) * m0;
  hipLaunchKernelGGL(sht::low_combine_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256),
                     0, s, *a);
  return hipGetLastError();
}

extern "C" hipError_t sht_low_allow_lds(size_t bytes) {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(sht::low_synthesis_kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(sht::low_analysis_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}
