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
  int blocks = a->job.nfields * a->mtiles * ((a->T + 1 + 63) / 64);
  if (blocks == 0) return hipSuccess;
  blocks = (blocks + 7) / 8 * 8;  // XCD-aware order (xcd_block)
  if (a->job.mode == sht::kVector)
    hipLaunchKernelGGL(sht::pack_kernel<2>, dim3(blocks), dim3(256), 0, s, *a);
  else
    hipLaunchKernelGGL(sht::pack_kernel<1>, dim3(blocks), dim3(256), 0, s, *a);
  return hipGetLastError();
}

extern "C" hipError_t sht_launch_unpack(const sht::UnpackArgs* a, hipStream_t s) {
  int blocks = a->job.nfields * a->tiles * ((a->T + 1 + 63) / 64);
  if (blocks == 0) return hipSuccess;
  blocks = (blocks + 7) / 8 * 8;  // XCD-aware order (xcd_block)
  if (a->job.vector)
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
