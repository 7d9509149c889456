// Spectral-side data movement: the boundary layout (*lead, T+1, T+1) complex64
// indexed [n, m] <-> packed per-order blocks C[c][m][p][n_idx] read and written
// by the grouped GEMM, plus the elementwise spectral operators.
//
// Both movers are tiled through LDS: a block owns 32 orders x 64 degrees of
// one field (or level), so each value is read once and global reads and
// writes are 128-256 byte runs on both sides.
//
// Packed values are fp32 in every precision mode; the split modes cut them
// into bf16 pieces as the GEMM loads them.
//
// Fused transforms (pack), per field or per vorticity/divergence level:
//   PLAIN   c, plus an optional per-field constant at (0, 0) (the reference's
//           add_grid_constant, transform.py:1832-1835, as lnps_reference.py:32-49
//           applies it); unpack adds one after an analysis likewise
//           (semilag/step.py:159-163)
//   IM      i*m*c                (reference transform.py:1752-1760)
//   ICOEF   i*c                  (zonal derivative against the P*m/cos table)
//   VECTOR  psi = zeta/lambda_n, chi = D/lambda_n, lambda_n = -n(n+1)/a^2,
//           n = 0 -> 0, m = 0 realified (reference transform.py:1804-1811),
//           written as the four columns [-psi, chi] per level.  The P*m/cos
//           segment needs [i chi, i psi] = [-a3, a2, a1, -a0] of those
//           columns; the GEMM reads that arrangement from the same block
//           through a column map instead of a second copy.
//   VECTOR_REC  the winds' orders m >= m0 in the recurrence form (vector jobs
//           whose orders m < m0 run in FP64): see pack_vector_rec_kernel.
#include "sht_kernels.h"

namespace sht {

constexpr int kTile = 32;   // orders per block
constexpr int kChunk = 64;  // degrees per pass of a block

// Block index of an XCD-aware launch (grid a multiple of 8): workgroups are
// dispatched round-robin over the 8 XCDs, so launch position b runs on XCD
// b % 8; consecutive logical blocks (neighbouring degree chunks of one field
// and order tile, which share the 128-byte lines where a packed run crosses a
// chunk edge) are given to one XCD, where they run side by side and the
// shared lines are fetched or written back once through that XCD's L2.
__device__ __forceinline__ int xcd_block() {
  const int q = (int)gridDim.x / 8;
  return (int)(blockIdx.x % 8) * q + (int)(blockIdx.x / 8);
}

// Degree of order m, parity p, in the chunk starting at n0, for lane j: the
// j-th degree of that parity in [n0, n0 + 64).
__device__ __forceinline__ int chunk_degree(int n0, int m, int p, int j) {
  return n0 + 2 * j + ((m + p - n0) & 1);
}

// The offsets koff[2m+p] of the tile's 32 orders x 2 parities, staged in LDS
// so that no data access waits on a dependent load.
__device__ __forceinline__ void stage_koff(int* skoff, const int* koff, int m0, int T) {
  if (threadIdx.x < 2 * kTile) {
    const int m = m0 + (int)(threadIdx.x >> 1);
    skoff[threadIdx.x] = m <= T ? koff[2 * m0 + threadIdx.x] : 0;
  }
}

// One block per (field or level, 32 orders, 64 degrees from the tile's first
// order): every input value is read once and every packed (m, p) run is
// written as 32 consecutive degrees.  Each thread issues all of its loads
// before it touches LDS, so a block keeps 16-32 KB in flight.
template <int NIN>
__global__ void __launch_bounds__(256) pack_kernel(PackArgs a) {
  __shared__ float2 tile[NIN][kChunk][kTile + 1];  // [input][n - n0][m - m0]
  __shared__ int skoff[2 * kTile];
  const int T = a.T, T1 = T + 1;
  const int nchunks = (T1 + kChunk - 1) / kChunk;
  const int b = xcd_block();
  const int c = b % nchunks;
  const int mt = (b / nchunks) % a.mtiles;
  const int f = b / nchunks / a.mtiles;
  if (f >= a.job.nfields) return;
  const int m0 = mt * kTile;
  const int n0 = m0 + kChunk * c;
  if (n0 > T) return;  // block-uniform
  constexpr int NCOL = 2 * NIN;
  constexpr int IT = kChunk * kTile / 256;  // loads per thread and input
  const size_t plane = (size_t)T1 * T1;
  const int tid = threadIdx.x;
  stage_koff(skoff, a.koff, m0, T);
  {
    float2 v[NIN][IT];
#pragma unroll
    for (int s = 0; s < NIN; ++s) {
      const float2* in = (s == 0 ? a.job.in0 : a.job.in1) + (size_t)f * plane;
#pragma unroll
      for (int it = 0; it < IT; ++it) {
        const int e = tid + 256 * it;
        const int n = n0 + e / kTile, m = m0 + e % kTile;
        const bool ok = n <= T && m <= n;  // m <= n <= T
        const float2 x = in[ok ? (size_t)n * T1 + m : 0];
        v[s][it] = ok ? x : make_float2(0.f, 0.f);
      }
    }
#pragma unroll
    for (int s = 0; s < NIN; ++s)
#pragma unroll
      for (int it = 0; it < IT; ++it) {
        const int e = tid + 256 * it;
        tile[s][e / kTile][e % kTile] = v[s][it];
      }
  }
  __syncthreads();
  const double a2 = (double)a.job.radius * (double)a.job.radius;
  // outputs (mi, p, j), j fastest: 32 consecutive packed degrees per run
  const int j = tid % 32, p = (tid / 32) % 2;
#pragma unroll
  for (int it = 0; it < kTile * 2 * 32 / 256; ++it) {
    const int mi = tid / 64 + 4 * it;
    const int m = m0 + mi;
    const int n = chunk_degree(n0, m, p, j);
    if (m > T || n > T || n < m + p) continue;
    const int nidx = (n - m - p) >> 1;
    const int r = n - n0;
    float o[4] = {0.f, 0.f, 0.f, 0.f};
    float2 cv = tile[0][r][mi];
    if (a.job.mode == kPlain) {
      if (a.job.off00 && n == 0) {  // m = 0 as well: (0, 0) is the only n = 0 entry
        const float2 d = a.job.off00[f];
        cv = make_float2(cv.x + d.x, cv.y + d.y);
      }
      o[0] = cv.x; o[1] = cv.y;
    } else if (a.job.mode == kIM) {
      o[0] = -(float)m * cv.y; o[1] = (float)m * cv.x;
    } else if (a.job.mode == kICoef) {
      o[0] = -cv.y; o[1] = cv.x;
    } else if (a.job.mode == kInvLap) {
      // the spectral inverse Laplacian's own arithmetic (n = 0 -> 0, m = 0 real)
      if (n >= 1) {
        const float eig = (float)(-((double)n * ((double)n + 1.0)) / a2);
        o[0] = cv.x / eig;
        o[1] = m == 0 ? 0.f : cv.y / eig;
      }
    } else if constexpr (NIN == 2) {
      const float2 d = tile[NIN - 1][r][mi];
      float2 psi = make_float2(0.f, 0.f), chi = make_float2(0.f, 0.f);
      if (n >= 1) {
        const float eig = (float)(-((double)n * ((double)n + 1.0)) / a2);
        psi = make_float2(cv.x / eig, cv.y / eig);
        chi = make_float2(d.x / eig, d.y / eig);
        if (m == 0) { psi.y = 0.f; chi.y = 0.f; }
      }
      o[0] = -psi.x; o[1] = -psi.y; o[2] = chi.x; o[3] = chi.y;
    }
    float* dst = a.job.out0 + ((size_t)a.job.col0 + (size_t)f * NCOL) * a.job.ldp +
                 skoff[2 * mi + p] + nidx;
#pragma unroll
    for (int cl = 0; cl < NCOL; ++cl) dst[(size_t)cl * a.job.ldp] = o[cl];
  }
}

// The recurrence form of the vector operators.  With the basis normalized as
// here, cos(phi) dP_n/dphi = (n+1) e_n P_{n-1} - n e_{n+1} P_{n+1}, where
// e_n = sqrt((n^2 - m^2) / (4 n^2 - 1)) (e_m = 0).  Substituted into
//   u = (-dpsi/dphi + (1/cos) dchi/dlambda) / a,
//   v = ((1/cos) dpsi/dlambda + dchi/dphi) / a,
// the winds of order m are contractions of the basis alone over the degrees
// k = m..T+1:
//   a cos(phi) u = sum_k P_k U_k,  U_k = (k-1) e_k psi_{k-1} - (k+2) e_{k+1} psi_{k+1} + i m chi_k
//   a cos(phi) v = sum_k P_k V_k,  V_k = i m psi_k - (k-1) e_k chi_{k-1} + (k+2) e_{k+1} chi_{k+1}
// (psi_k = 0 outside m <= k <= T), one plane contraction per wind component
// against the P*m/cos and dP/dphi tables' two.  The pack forms U and V in FP64
// from psi, chi = zeta / lambda_n, D / lambda_n (fp32, as the other vector
// pack forms them) and writes them in fp32 as the four
// columns [U, V] per level; the unfold multiplies by 1/(a cos(phi)).  Near a
// pole the terms of order 0 cancel to O(cos^2), so the form is used only where
// the orders below m0 >= 4 run in FP64, or from m = mdp = 4 on, the orders
// below in the two-segment form's columns [-psi, chi] (the plain FP32 path);
// from there every term carries cos^m.
// A block owns 32 orders x 64 degrees plus the degrees either side (halo rows
// n0 - 1 and n0 + 64), so neighbours in degree come from LDS.
// e_n^m = sqrt((n^2 - m^2) / (4 n^2 - 1)) in FP64 without a divide or a
// square root: x rsqrt(x y) from an fp32 seed refined by two Newton steps
// r <- r (3 - z r^2) / 2 (relative error ~1e-16; x y is exact); 0 for n <= m.
__device__ __forceinline__ double rec_eps(int n, int m) {
  const double x = (double)n * n - (double)m * m, y = 4.0 * (double)n * n - 1.0;
  if (!(x > 0.0)) return 0.0;
  const double z = x * y;
  double r = (double)__builtin_amdgcn_rsqf((float)z);
  r = r * (1.5 - 0.5 * z * r * r);
  r = r * (1.5 - 0.5 * z * r * r);
  return x * r;
}

// 1 / lambda_n = -a^2 / (n (n+1)) in FP64 without a divide: an fp32
// reciprocal seed refined by two Newton steps r <- r (2 - d r); 0 at n = 0.
__device__ __forceinline__ double inv_lambda(double a2, int n) {
  if (n < 1) return 0.0;
  const double d = (double)n * ((double)n + 1.0);
  double r = (double)__builtin_amdgcn_rcpf((float)d);
  r = r * (2.0 - d * r);
  r = r * (2.0 - d * r);
  return -a2 * r;
}

__global__ void __launch_bounds__(256) pack_vector_rec_kernel(PackArgs a) {
  constexpr int kRows = kChunk + 2;
  __shared__ float2 tile[2][kRows][kTile + 1];  // [input][n - n0 + 1][m - m0]
  __shared__ int skoff[2 * kTile];
  const int T = a.T, T1 = T + 1;
  const int nchunks = (T + 2 + kChunk - 1) / kChunk;  // degrees up to T+1
  const int b = xcd_block();
  const int c = b % nchunks;
  const int mt = (b / nchunks) % a.mtiles;
  const int f = b / nchunks / a.mtiles;
  if (f >= a.job.nfields) return;
  const int m0 = mt * kTile;
  const int n0 = m0 + kChunk * c;
  if (n0 > T + 1 || m0 + kTile <= a.job.m0) return;  // block-uniform
  constexpr int E = kRows * kTile;
  constexpr int IT = (E + 255) / 256;
  const size_t plane = (size_t)T1 * T1;
  const int tid = threadIdx.x;
  stage_koff(skoff, a.koff, m0, T);
  {
    float2 v[2][IT];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const float2* in = (s == 0 ? a.job.in0 : a.job.in1) + (size_t)f * plane;
#pragma unroll
      for (int it = 0; it < IT; ++it) {
        const int e = tid + 256 * it;
        const int n = n0 - 1 + e / kTile, m = m0 + e % kTile;
        const bool ok = e < E && n <= T && m <= n;  // m <= n <= T
        const float2 x = in[ok ? (size_t)n * T1 + m : 0];
        v[s][it] = ok ? x : make_float2(0.f, 0.f);
      }
    }
    const double a2 = (double)a.job.radius * (double)a.job.radius;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = tid + 256 * it;
      // the tile holds psi, chi in fp32 (FP64 here, between the loads and the
      // LDS stores, costs a quarter of the kernel's time)
      const float g = (float)inv_lambda(a2, n0 - 1 + e / kTile);
#pragma unroll
      for (int s = 0; s < 2; ++s)
        if (e < E) tile[s][e / kTile][e % kTile] = make_float2(g * v[s][it].x, g * v[s][it].y);
    }
  }
  __syncthreads();
  const int j = tid % 32, p = (tid / 32) % 2;
#pragma unroll
  for (int it = 0; it < kTile * 2 * 32 / 256; ++it) {
    const int mi = tid / 64 + 4 * it;
    const int m = m0 + mi;
    const int n = chunk_degree(n0, m, p, j);
    if (m > T || m < a.job.m0 || n > T + 1 || n < m + p) continue;
    const int nidx = (n - m - p) >> 1;
    const int r = n - n0 + 1;
    // psi, chi of degree n + d from tile row r + d (zero outside m <= deg <= T)
    auto pot = [&](int s, int d) {
      const float2 z = tile[s][r + d][mi];
      return make_double2((double)z.x, (double)z.y);
    };
    const double2 pm = pot(0, -1), p0 = pot(0, 0), pp = pot(0, 1);
    const double2 cm = pot(1, -1), c0 = pot(1, 0), cp = pot(1, 1);
    const double dn = (double)n, dm = (double)m;
    float o[4];
    if (m < a.job.mdp) {  // the two-segment form's [-psi, chi] (no degree T+1)
      if (n > T) continue;
      const float2 ps = tile[0][r][mi], ch = tile[1][r][mi];
      o[0] = -ps.x;
      o[1] = m == 0 ? 0.f : -ps.y;
      o[2] = ch.x;
      o[3] = m == 0 ? 0.f : ch.y;
    } else {
      const double ca = (dn - 1.0) * rec_eps(n, m), cb = (dn + 2.0) * rec_eps(n + 1, m);
      o[0] = (float)(ca * pm.x - cb * pp.x - dm * c0.y);  // U
      o[1] = (float)(ca * pm.y - cb * pp.y + dm * c0.x);
      o[2] = (float)(-dm * p0.y - ca * cm.x + cb * cp.x);  // V
      o[3] = (float)(dm * p0.x - ca * cm.y + cb * cp.y);
    }
    float* dst = a.job.out0 + ((size_t)a.job.col0 + (size_t)f * 4) * a.job.ldp +
                 skoff[2 * mi + p] + nidx;
#pragma unroll
    for (int cl = 0; cl < 4; ++cl) dst[(size_t)cl * a.job.ldp] = o[cl];
  }
}

// One block per (field or level, 32 orders, 64 rows).  Entries above the
// triangle (m > n) are written as zeros; chunks that lie wholly above it need
// no reads.  As in pack, every thread issues all of its loads first.
template <int NCOL>
__global__ void __launch_bounds__(256) unpack_kernel(UnpackArgs a) {
  __shared__ float tile[NCOL][kChunk][kTile + 1];  // [col][n - n0][m - m0]
  __shared__ int skoff[2 * kTile];
  const int T = a.T, T1 = T + 1;
  const int nchunks = (T1 + kChunk - 1) / kChunk;
  const int b = xcd_block();
  const int c = b % nchunks;
  const int mt = (b / nchunks) % a.tiles;
  const int f = b / nchunks / a.tiles;
  if (f >= a.job.nfields) return;
  const int m0 = mt * kTile;
  const int n0 = kChunk * c;
  constexpr int NOUT = NCOL / 2;
  constexpr int IT = kTile * 2 * 32 / 256;  // loads per thread and column
  const size_t plane = (size_t)T1 * T1;
  const int tid = threadIdx.x;
  if (n0 + kChunk - 1 >= m0) {  // some m <= n in this chunk (block-uniform)
    stage_koff(skoff, a.koff, m0, T);
    __syncthreads();
    // gather each (col, m, p) run of packed degrees that falls in the chunk
    const int j = tid % 32, p = (tid / 32) % 2;
    const float* in = a.job.in + ((size_t)a.job.col0 + (size_t)f * NCOL) * a.job.ldp;
    float v[NCOL][IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int mi = tid / 64 + 4 * it;
      const int m = m0 + mi;
      const int n = chunk_degree(n0, m, p, j);
      const bool ok = m <= T && n <= T && n >= m + p;
      const size_t idx = ok ? (size_t)skoff[2 * mi + p] + ((n - m - p) >> 1) : 0;
#pragma unroll
      for (int cl = 0; cl < NCOL; ++cl) {
        const float x = in[(size_t)cl * a.job.ldp + idx];
        v[cl][it] = (ok && !(m == 0 && (cl & 1))) ? x : 0.f;  // m = 0 realified
      }
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int mi = tid / 64 + 4 * it;
      const int n = chunk_degree(n0, m0 + mi, p, j);
      if (n - n0 < kChunk) {
#pragma unroll
        for (int cl = 0; cl < NCOL; ++cl) tile[cl][n - n0][mi] = v[cl][it];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int o = 0; o < NOUT; ++o) {
    float2* out = (o == 0 ? a.job.out0 : a.job.out1) + (size_t)f * plane;
#pragma unroll
    for (int it = 0; it < kChunk * kTile / 256; ++it) {
      const int e = tid + 256 * it;
      const int mi = e % kTile, r = e / kTile;
      const int n = n0 + r, m = m0 + mi;
      if (n > T || m > T) continue;
      if (m < a.job.m0 && m <= n) continue;  // the FP64 low-order kernel's entries
      float2 v = m <= n ? make_float2(tile[2 * o][r][mi], tile[2 * o + 1][r][mi])
                        : make_float2(0.f, 0.f);
      if (NCOL == 2 && a.job.off00 && n == 0 && m == 0) {
        const float2 d = a.job.off00[f];
        v = make_float2(v.x + d.x, v.y + d.y);
      }
      out[(size_t)n * T1 + m] = v;
    }
  }
}

// The analysis side of the recurrence form (orders m >= m0 of vector jobs
// whose orders below m0 run in FP64).  The fold weights the winds by
// 2 pi w / (a nlon cos(phi)) and the GEMM contracts the basis over degrees
// m..T+1, giving [iB, iA] per level: A_k = sum_j (2 pi w_j / (a cos_j)) u_j P_k,
// B_k likewise of v.  With the derivative recurrence above,
//   D_n    = i m A_n + n e_{n+1} B_{n+1} - (n+1) e_n B_{n-1}
//   zeta_n = i m B_n - n e_{n+1} A_{n+1} + (n+1) e_n A_{n-1},
// formed here in FP64 from the packed columns (degrees n - 1 .. n + 1 of the
// chunk's rows and its halo).  The orders m < m0 are left to the FP64
// low-order kernels (which form them the same way in FP64), the orders
// m < mdp hold [zeta, D] from the two-segment form; entries above the
// triangle are written as zeros.
__global__ void __launch_bounds__(256) unpack_vector_rec_kernel(UnpackArgs a) {
  constexpr int kRows = kChunk + 2;
  __shared__ float tile[4][kRows][kTile + 1];  // [col][n - n0 + 1][m - m0]
  __shared__ int skoff[2 * kTile];
  const int T = a.T, T1 = T + 1;
  const int nchunks = (T1 + kChunk - 1) / kChunk;
  const int b = xcd_block();
  const int c = b % nchunks;
  const int mt = (b / nchunks) % a.tiles;
  const int f = b / nchunks / a.tiles;
  if (f >= a.job.nfields) return;
  const int m0 = mt * kTile;
  const int n0 = kChunk * c;
  constexpr int IT = kTile * 2 * 32 / 256;  // loads per thread and column
  const size_t plane = (size_t)T1 * T1;
  const int tid = threadIdx.x;
  const float* in = a.job.in + ((size_t)a.job.col0 + (size_t)f * 4) * a.job.ldp;
  float2* out0 = a.job.out0 + (size_t)f * plane;
  float2* out1 = a.job.out1 + (size_t)f * plane;
  if (n0 + kChunk >= m0) {  // some m <= n + 1 in this chunk (block-uniform)
    stage_koff(skoff, a.koff, m0, T);
    __syncthreads();
    const int j = tid % 32, p = (tid / 32) % 2;
    float v[4][IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int mi = tid / 64 + 4 * it;
      const int m = m0 + mi;
      const int n = chunk_degree(n0, m, p, j);
      const bool ok = m <= T && n <= T + 1 && n >= m + p;
      const size_t idx = ok ? (size_t)skoff[2 * mi + p] + ((n - m - p) >> 1) : 0;
#pragma unroll
      for (int cl = 0; cl < 4; ++cl) {
        const float x = in[(size_t)cl * a.job.ldp + idx];
        v[cl][it] = ok ? x : 0.f;
      }
    }
    // halo: degrees n0 - 1 and n0 + 64 of the tile's orders
    float h[4] = {0.f, 0.f, 0.f, 0.f};
    const int hm = m0 + (tid & 31), hn = (tid >> 5) ? n0 + kChunk : n0 - 1;
    if (tid < 64) {
      const bool ok = hm <= T && hn >= hm && hn <= T + 1;
      const int hp = (hn - hm) & 1;
      const size_t idx = ok ? (size_t)skoff[2 * (tid & 31) + hp] + ((hn - hm - hp) >> 1) : 0;
#pragma unroll
      for (int cl = 0; cl < 4; ++cl) {
        const float x = in[(size_t)cl * a.job.ldp + idx];
        h[cl] = ok ? x : 0.f;
      }
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int mi = tid / 64 + 4 * it;
      const int n = chunk_degree(n0, m0 + mi, p, j);
      if (n - n0 < kChunk) {
#pragma unroll
        for (int cl = 0; cl < 4; ++cl) tile[cl][n - n0 + 1][mi] = v[cl][it];
      }
    }
    if (tid < 64) {
#pragma unroll
      for (int cl = 0; cl < 4; ++cl) tile[cl][(tid >> 5) ? kRows - 1 : 0][tid & 31] = h[cl];
    }
    __syncthreads();
  }
#pragma unroll
  for (int it = 0; it < kChunk * kTile / 256; ++it) {
    const int e = tid + 256 * it;
    const int mi = e % kTile, r = e / kTile;
    const int n = n0 + r, m = m0 + mi;
    if (n > T || m > T) continue;
    if (m < a.job.m0 && m <= n) continue;  // the FP64 low-order kernels' entries
    float2 z = make_float2(0.f, 0.f), d = make_float2(0.f, 0.f);
    if (m < a.job.mdp && m <= n) {  // the two-segment form: [zeta, D] as they are
      const int rr = r + 1;
      z = make_float2(tile[0][rr][mi], m == 0 ? 0.f : tile[1][rr][mi]);
      d = make_float2(tile[2][rr][mi], m == 0 ? 0.f : tile[3][rr][mi]);
    } else if (m <= n) {
      const int rr = r + 1;
      const double dn = (double)n, dm = (double)m;
      const double ca = (dn + 1.0) * rec_eps(n, m), cb = dn * rec_eps(n + 1, m);  // X_{n-1}, X_{n+1}
      // columns 0, 1: i B; 2, 3: i A
      const double bx = tile[0][rr][mi], by = tile[1][rr][mi];
      const double ax = tile[2][rr][mi], ay = tile[3][rr][mi];
      const double sax = cb * tile[2][rr + 1][mi] - ca * tile[2][rr - 1][mi];
      const double say = cb * tile[3][rr + 1][mi] - ca * tile[3][rr - 1][mi];
      const double sbx = cb * tile[0][rr + 1][mi] - ca * tile[0][rr - 1][mi];
      const double sby = cb * tile[1][rr + 1][mi] - ca * tile[1][rr - 1][mi];
      // zeta = m (iB) + i (cb iA_{n+1} - ca iA_{n-1}); D = m (iA) - i (cb iB_{n+1} - ca iB_{n-1})
      z = make_float2((float)(dm * bx - say), (float)(dm * by + sax));
      d = make_float2((float)(dm * ax + sby), (float)(dm * ay - sbx));
      if (m == 0) {
        z.y = 0.f;
        d.y = 0.f;
      }
    }
    out0[(size_t)n * T1 + m] = z;
    out1[(size_t)n * T1 + m] = d;
  }
}

// Elementwise spectral operators on the boundary layout.
//   op 0: project   (zero above the triangle, m = 0 realified; transform.py:1124-1144)
//   op 1: laplacian (c * lambda_n; transform.py:1798-1802)
//   op 2: inverse laplacian (c / lambda_n for n >= 1, n = 0 -> 0, projected;
//         transform.py:1804-1811)
__global__ void spectral_elementwise_kernel(const float2* __restrict__ in,
                                            float2* __restrict__ out, long long total, int T,
                                            int op, float radius) {
  const int T1 = T + 1;
  const double a2 = (double)radius * (double)radius;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int m = (int)(e % T1);
    const int n = (int)((e / T1) % T1);
    float2 c = in[e];
    const float eig = (float)(-((double)n * ((double)n + 1.0)) / a2);
    float2 r;
    if (op == 1) {
      r = make_float2(c.x * eig, c.y * eig);
    } else {
      if (op == 2) c = (n >= 1) ? make_float2(c.x / eig, c.y / eig) : make_float2(0.f, 0.f);
      if (m > n) r = make_float2(0.f, 0.f);
      else if (m == 0) r = make_float2(c.x, 0.f);
      else r = c;
    }
    out[e] = r;
  }
}

}  // namespace sht

extern "C" hipError_t sht_launch_pack(const sht::PackArgs* a, hipStream_t s) {
  const int degrees = a->job.mode == sht::kVectorRec ? a->T + 2 : a->T + 1;
  int blocks = a->job.nfields * a->mtiles * ((degrees + 63) / 64);
  if (blocks == 0) return hipSuccess;
  blocks = (blocks + 7) / 8 * 8;  // XCD-aware order (xcd_block)
  if (a->job.mode == sht::kVectorRec)
    hipLaunchKernelGGL(sht::pack_vector_rec_kernel, dim3(blocks), dim3(256), 0, s, *a);
  else if (a->job.mode == sht::kVector)
    hipLaunchKernelGGL(sht::pack_kernel<2>, dim3(blocks), dim3(256), 0, s, *a);
  else
    hipLaunchKernelGGL(sht::pack_kernel<1>, dim3(blocks), dim3(256), 0, s, *a);
  return hipGetLastError();
}

extern "C" hipError_t sht_launch_unpack(const sht::UnpackArgs* a, hipStream_t s) {
  int blocks = a->job.nfields * a->tiles * ((a->T + 1 + 63) / 64);
  if (blocks == 0) return hipSuccess;
  blocks = (blocks + 7) / 8 * 8;  // XCD-aware order (xcd_block)
  if (a->job.rec)
    hipLaunchKernelGGL(sht::unpack_vector_rec_kernel, dim3(blocks), dim3(256), 0, s, *a);
  else if (a->job.vector)
    hipLaunchKernelGGL(sht::unpack_kernel<4>, dim3(blocks), dim3(256), 0, s, *a);
  else
    hipLaunchKernelGGL(sht::unpack_kernel<2>, dim3(blocks), dim3(256), 0, s, *a);
  return hipGetLastError();
}

extern "C" hipError_t sht_launch_spectral_elementwise(const float2* in, float2* out,
                                                      long long total, int T, int op,
                                                      float radius, hipStream_t s) {
  if (total <= 0) return hipSuccess;
  long long blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(sht::spectral_elementwise_kernel, dim3((unsigned)blocks), dim3(256), 0, s,
                     in, out, total, T, op, radius);
  return hipGetLastError();
}
