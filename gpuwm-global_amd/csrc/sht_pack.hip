// Spectral-side data movement: the boundary layout (*lead, T+1, T+1) complex64
// indexed [n, m] <-> packed per-order blocks C[c][m][p][n_idx] read and written
// by the grouped GEMM, plus the elementwise spectral operators.
//
// Both movers are tiled through LDS: a block owns 32 orders x 32 packed
// degrees (pack) or 32 degrees x 32 orders (unpack) of one field, so global
// reads and writes are 64-256 byte runs on both sides.
//
// For the split-precision plans the packed values are written as bf16 pieces
// (piece planes `pstride` elements apart), the form the GEMM reads.
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






constexpr int kTile = 32;
constexpr int kRows = 3 * kTile - 1;  // input degrees spanned by one pack tile

__global__ void __launch_bounds__(256) pack_kernel(PackArgs a) {
  __shared__ float2 tile[2][kRows][kTile + 1];
  const int T = a.T, T1 = T + 1;
  int b = blockIdx.x;
  const int nt = b % a.ntiles;
  b /= a.ntiles;
  const int mt = b % a.mtiles;
  const int f = b / a.mtiles;
  if (f >= a.job.nfields) return;
  const int m0 = mt * kTile, t0 = nt * kTile;
  const int nlo = m0 + 2 * t0;
  const bool vec = a.job.mode == kVector;
  const int nin = vec ? 2 : 1;
  const size_t plane = (size_t)T1 * T1;
  for (int s = 0; s < nin; ++s) {
    const float2* in = (s == 0 ? a.job.in0 : a.job.in1) + (size_t)f * plane;
    for (int e = threadIdx.x; e < kRows * kTile; e += blockDim.x) {
      const int r = e / kTile, mi = e % kTile;
      const int n = nlo + r, m = m0 + mi;
      float2 v = make_float2(0.f, 0.f);
      if (n <= T && m <= T && m <= n) v = in[(size_t)n * T1 + m];
      tile[s][r][mi] = v;
    }
  }
  __syncthreads();
  const double a2 = (double)a.job.radius * (double)a.job.radius;
  const int ncol = vec ? 4 : 2;
  // outputs: (mi, p, tj) with tj fastest
  for (int e = threadIdx.x; e < kTile * 2 * kTile; e += blockDim.x) {
    const int tj = e % kTile;
    const int p = (e / kTile) % 2;
    const int mi = e / (2 * kTile);
    const int m = m0 + mi;
    if (m > T) continue;
    const int nidx = t0 + tj;
    const int kbeg = a.koff[2 * m + p];
    const int kp = a.koff[2 * m + p + 1] - kbeg;
    if (nidx >= kp) continue;
    const int n = m + p + 2 * nidx;
    const int r = mi + p + 2 * tj;  // row in the tile
    float o[4] = {0.f, 0.f, 0.f, 0.f};
    if (n <= T) {
      float2 c = tile[0][r][mi];
      if (a.job.mode == kPlain) {
        o[0] = c.x; o[1] = c.y;
      } else if (a.job.mode == kIM) {
        o[0] = -(float)m * c.y; o[1] = (float)m * c.x;
      } else if (a.job.mode == kICoef) {
        o[0] = -c.y; o[1] = c.x;
      } else {
        float2 d = tile[1][r][mi];
        float2 psi = make_float2(0.f, 0.f), chi = make_float2(0.f, 0.f);
        if (n >= 1) {
          const float eig = (float)(-((double)n * ((double)n + 1.0)) / a2);
          psi = make_float2(c.x / eig, c.y / eig);
          chi = make_float2(d.x / eig, d.y / eig);
          if (m == 0) { psi.y = 0.f; chi.y = 0.f; }
        }
        o[0] = -psi.x; o[1] = -psi.y; o[2] = chi.x; o[3] = chi.y;
      }
    }
    for (int cl = 0; cl < ncol; ++cl) {
      const size_t col = (size_t)a.job.col0 + (size_t)f * ncol + cl;
      const size_t idx = col * a.job.ldp + kbeg + nidx;
      if (a.job.split == 0)
        a.job.out0[idx] = o[cl];
      else
        store_bf16_pieces(reinterpret_cast<unsigned short*>(a.job.out0) + idx, a.job.pstride,
                          a.job.split, o[cl]);
    }
  }
}





__global__ void __launch_bounds__(256) unpack_kernel(UnpackArgs a) {
  __shared__ float tile[4][kTile][kTile + 1];  // [col][n - n0][m - m0]
  const int T = a.T, T1 = T + 1;
  int b = blockIdx.x;
  const int mt = b % a.tiles;
  b /= a.tiles;
  const int ntl = b % a.tiles;
  const int f = b / a.tiles;
  if (f >= a.job.nfields) return;
  const int n0 = ntl * kTile, m0 = mt * kTile;
  const int ncol = a.job.vector ? 4 : 2;
  // Zero the tile, then gather each (m, p) run of packed degrees.
  for (int e = threadIdx.x; e < 4 * kTile * (kTile + 1); e += blockDim.x)
    (&tile[0][0][0])[e] = 0.f;
  __syncthreads();
  for (int e = threadIdx.x; e < ncol * kTile * 2 * 16; e += blockDim.x) {
    const int j = e % 16;
    const int p = (e / 16) % 2;
    const int mi = (e / 32) % kTile;
    const int cl = e / (32 * kTile);
    const int m = m0 + mi;
    if (m > T) continue;
    // first packed index whose degree is >= n0
    int base = n0 - m - p;
    int jlo = base <= 0 ? 0 : (base + 1) / 2;
    const int nidx = jlo + j;
    const int n = m + p + 2 * nidx;
    if (n < n0 || n >= n0 + kTile || n > T) continue;
    const int kbeg = a.koff[2 * m + p];
    const size_t col = (size_t)a.job.col0 + (size_t)f * ncol + cl;
    float v = a.job.in[col * a.job.ldp + kbeg + nidx];
    if (m == 0 && (cl & 1)) v = 0.f;  // m = 0 realified
    tile[cl][n - n0][mi] = v;
  }
  __syncthreads();
  const size_t plane = (size_t)T1 * T1;
  const int nout = a.job.vector ? 2 : 1;
  for (int o = 0; o < nout; ++o) {
    float2* out = (o == 0 ? a.job.out0 : a.job.out1) + (size_t)f * plane;
    for (int e = threadIdx.x; e < kTile * kTile; e += blockDim.x) {
      const int mi = e % kTile, ni = e / kTile;
      const int n = n0 + ni, m = m0 + mi;
      if (n > T || m > T) continue;
      out[(size_t)n * T1 + m] = make_float2(tile[2 * o][ni][mi], tile[2 * o + 1][ni][mi]);
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
  int blocks = a->job.nfields * a->mtiles * a->ntiles;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(sht::pack_kernel, dim3(blocks), dim3(256), 0, s, *a);
  return hipGetLastError();
}

extern "C" hipError_t sht_launch_unpack(const sht::UnpackArgs* a, hipStream_t s) {
  int blocks = a->job.nfields * a->tiles * a->tiles;
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
