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
//   PLAIN   c
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

// Degree of order m, parity p, in the chunk starting at n0, for lane j: the
// j-th degree of that parity in [n0, n0 + 64).
__device__ __forceinline__ int chunk_degree(int n0, int m, int p, int j) {
  return n0 + 2 * j + ((m + p - n0) & 1);
}

// One block per (field or level, 32 orders, 64 degrees from the tile's first
// order): every input value is read once and every packed (m, p) run is
// written as 32 consecutive degrees.
__global__ void __launch_bounds__(256) pack_kernel(PackArgs a) {
  __shared__ float2 tile[2][kChunk][kTile + 1];  // [input][n - n0][m - m0]
  const int T = a.T, T1 = T + 1;
  const int nchunks = (T1 + kChunk - 1) / kChunk;
  const int c = blockIdx.x % nchunks;
  const int mt = (blockIdx.x / nchunks) % a.mtiles;
  const int f = blockIdx.x / nchunks / a.mtiles;
  if (f >= a.job.nfields) return;
  const int m0 = mt * kTile;
  const int n0 = m0 + kChunk * c;
  if (n0 > T) return;  // block-uniform
  const bool vec = a.job.mode == kVector;
  const int nin = vec ? 2 : 1;
  const int ncol = vec ? 4 : 2;
  const size_t plane = (size_t)T1 * T1;
  const double a2 = (double)a.job.radius * (double)a.job.radius;
  {
    for (int s = 0; s < nin; ++s) {
      const float2* in = (s == 0 ? a.job.in0 : a.job.in1) + (size_t)f * plane;
      for (int e = threadIdx.x; e < kChunk * kTile; e += blockDim.x) {
        const int r = e / kTile, mi = e % kTile;
        const int n = n0 + r, m = m0 + mi;
        float2 v = make_float2(0.f, 0.f);
        if (n <= T && m <= T && m <= n) v = in[(size_t)n * T1 + m];
        tile[s][r][mi] = v;
      }
    }
    __syncthreads();
    // outputs (mi, p, j), j fastest: 32 consecutive packed degrees per run
    for (int e = threadIdx.x; e < kTile * 2 * 32; e += blockDim.x) {
      const int j = e % 32;
      const int p = (e / 32) % 2;
      const int mi = e / 64;
      const int m = m0 + mi;
      if (m > T) continue;
      const int n = chunk_degree(n0, m, p, j);
      if (n > T || n < m + p) continue;
      const int nidx = (n - m - p) >> 1;
      const int kbeg = a.koff[2 * m + p];
      const int r = n - n0;
      float o[4] = {0.f, 0.f, 0.f, 0.f};
      const float2 c = tile[0][r][mi];
      if (a.job.mode == kPlain) {
        o[0] = c.x; o[1] = c.y;
      } else if (a.job.mode == kIM) {
        o[0] = -(float)m * c.y; o[1] = (float)m * c.x;
      } else if (a.job.mode == kICoef) {
        o[0] = -c.y; o[1] = c.x;
      } else {
        const float2 d = tile[1][r][mi];
        float2 psi = make_float2(0.f, 0.f), chi = make_float2(0.f, 0.f);
        if (n >= 1) {
          const float eig = (float)(-((double)n * ((double)n + 1.0)) / a2);
          psi = make_float2(c.x / eig, c.y / eig);
          chi = make_float2(d.x / eig, d.y / eig);
          if (m == 0) { psi.y = 0.f; chi.y = 0.f; }
        }
        o[0] = -psi.x; o[1] = -psi.y; o[2] = chi.x; o[3] = chi.y;
      }
      for (int cl = 0; cl < ncol; ++cl) {
        const size_t col = (size_t)a.job.col0 + (size_t)f * ncol + cl;
        a.job.out0[col * a.job.ldp + kbeg + nidx] = o[cl];
      }
    }
  }
}

// One block per (field or level, 32 orders, 64 rows).  Entries above the
// triangle (m > n) are written as zeros; chunks that lie wholly above it need
// no reads.
__global__ void __launch_bounds__(256) unpack_kernel(UnpackArgs a) {
  __shared__ float tile[4][kChunk][kTile + 1];  // [col][n - n0][m - m0]
  const int T = a.T, T1 = T + 1;
  const int nchunks = (T1 + kChunk - 1) / kChunk;
  const int c = blockIdx.x % nchunks;
  const int mt = (blockIdx.x / nchunks) % a.tiles;
  const int f = blockIdx.x / nchunks / a.tiles;
  if (f >= a.job.nfields) return;
  const int m0 = mt * kTile;
  const int n0 = kChunk * c;
  const int ncol = a.job.vector ? 4 : 2;
  const int nout = a.job.vector ? 2 : 1;
  const size_t plane = (size_t)T1 * T1;
  {
    const bool any = n0 + kChunk - 1 >= m0;  // some m <= n in this chunk
    if (any) {
      // gather each (col, m, p) run of packed degrees that falls in the chunk
      for (int e = threadIdx.x; e < ncol * kTile * 2 * 32; e += blockDim.x) {
        const int j = e % 32;
        const int p = (e / 32) % 2;
        const int mi = (e / 64) % kTile;
        const int cl = e / (64 * kTile);
        const int m = m0 + mi;
        if (m > T) continue;
        const int n = chunk_degree(n0, m, p, j);
        if (n > T || n < m + p) continue;
        const int nidx = (n - m - p) >> 1;
        const size_t col = (size_t)a.job.col0 + (size_t)f * ncol + cl;
        float v = a.job.in[col * a.job.ldp + a.koff[2 * m + p] + nidx];
        if (m == 0 && (cl & 1)) v = 0.f;  // m = 0 realified
        tile[cl][n - n0][mi] = v;
      }
      __syncthreads();
    }
    for (int o = 0; o < nout; ++o) {
      float2* out = (o == 0 ? a.job.out0 : a.job.out1) + (size_t)f * plane;
      for (int e = threadIdx.x; e < kChunk * kTile; e += blockDim.x) {
        const int mi = e % kTile, r = e / kTile;
        const int n = n0 + r, m = m0 + mi;
        if (n > T || m > T) continue;
        const float2 v = m <= n ? make_float2(tile[2 * o][r][mi], tile[2 * o + 1][r][mi])
                                : make_float2(0.f, 0.f);
        out[(size_t)n * T1 + m] = v;
      }
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
  hipLaunchKernelGGL(sht::pack_kernel, dim3(blocks), dim3(256), 0, s, *a);
  return hipGetLastError();
}

extern "C" hipError_t sht_launch_unpack(const sht::UnpackArgs* a, hipStream_t s) {
  int blocks = a->job.nfields * a->tiles * ((a->T + 1 + 63) / 64);
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(sht::unpack_kernel, dim3(blocks), dim3(256), 0, s, *a);
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
