// Grouped Legendre GEMM over the triangular m-schedule (one persistent launch).
//
// Every Legendre contraction of a batch -- every job, order m, output parity
// and tile -- is one entry of a work list; each workgroup walks the list
// with a grid stride, so one launch covers all fields, levels, members and
// truncations of the batch.  A problem is
//     C[r][c] = sum_seg sum_k A_seg[r][k] * B_seg[k][c]
// with A a Legendre table block (k contiguous, rows and k zero padded to 32)
// and B a data block (fp32, arbitrary k and c strides).  Two segments
// express the vector operators (P*m/cos and dP/dphi against differently
// arranged data) as one accumulation.
//
// Precision modes (template):
//   FP32   v_mfma_f32_16x16x4_f32, fp32 tables and data.
//   BF16X3 v_mfma_f32_16x16x32_bf16 on hi+lo pieces: a0b0 + a0b1 + a1b0.
//   BF16X6 three pieces each: a0b0 + (a0b1 + a1b0) + (a0b2 + a1b1 + a2b0),
//          every term down to 2^-16 relative; the five small products
//          accumulate apart from a0b0 and are added once at the end.
// Tables are pre-split on the device from FP64 (sht_tables.hip); data are
// split here, in registers, by round-to-nearest-even conversion.
//
// Tile: 256 threads = 4 waves, block tile 64 rows x 64 columns, each wave a
// 32x32 quadrant as 2x2 MFMA 16x16 tiles.  Lane l = (l & 15) row/column,
// (l >> 4) selects a run of 8 consecutive k; A and B fragments use the same
// k mapping, so each MFMA sums matching k.
#include "sht_kernels.h"

namespace sht {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) unsigned short u16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;





struct GemmWork {
  const GemmProb* probs;
  const int4* tiles;    // {prob, r0, c0, unused}
  int ntiles;
};

__device__ __forceinline__ unsigned short f2bf(float f) {
  __bf16 h = (__bf16)f;  // round to nearest even
  return __builtin_bit_cast(unsigned short, h);
}
__device__ __forceinline__ float bf2f(unsigned short b) {
  return __uint_as_float(((unsigned int)b) << 16);
}

template <int NP>
__device__ __forceinline__ void split_pieces(const float (&x)[8], u16x8 (&out)[3]) {
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    float r = x[s];
    unsigned short b0 = f2bf(r);
    out[0][s] = b0;
    if constexpr (NP >= 2) {
      r = r - bf2f(b0);
      unsigned short b1 = f2bf(r);
      out[1][s] = b1;
      if constexpr (NP >= 3) {
        r = r - bf2f(b1);
        out[2][s] = f2bf(r);
      }
    }
  }
}

__device__ __forceinline__ void load_b8(const float* __restrict__ B, int bks, float (&x)[8]) {
  if (bks == 1) {
    const f32x4* p = reinterpret_cast<const f32x4*>(B);
    f32x4 u = p[0], v = p[1];
    x[0] = u[0]; x[1] = u[1]; x[2] = u[2]; x[3] = u[3];
    x[4] = v[0]; x[5] = v[1]; x[6] = v[2]; x[7] = v[3];
  } else {
#pragma unroll
    for (int s = 0; s < 8; ++s) x[s] = B[(long long)s * bks];
  }
}

template <int MODE>
__global__ void __launch_bounds__(256) legendre_gemm_kernel(GemmWork w) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int li = lane & 15;
  const int kg = lane >> 4;  // k run: k0 + 8*kg .. +7
  constexpr int NP = MODE == kBF16X6 ? 3 : 2;

  for (int t = blockIdx.x; t < w.ntiles; t += gridDim.x) {
    const int4 tile = w.tiles[t];
    const GemmProb& P = w.probs[tile.x];
    const int rbase = tile.y + 32 * (wave >> 1);
    const int cbase = tile.z + 32 * (wave & 1);
    if (rbase >= P.R || cbase >= P.N) continue;  // wave-uniform

    f32x4 acc[2][2], acc2[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
        acc2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
      }

    for (int sgi = 0; sgi < P.nseg; ++sgi) {
      const GemmSeg S = P.seg[sgi];
      for (int k0 = 0; k0 < S.K; k0 += 32) {
        const int kk = k0 + 8 * kg;
        // B fragments (two 16-column blocks), fp32.
        float bx[2][8];
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
          const int col = cbase + 16 * cb + li;
          if (col < P.N) {
            load_b8(S.B + (long long)kk * S.bks + (long long)col * S.bcs, S.bks, bx[cb]);
          } else {
#pragma unroll
            for (int s = 0; s < 8; ++s) bx[cb][s] = 0.f;
          }
        }
        if constexpr (MODE == kFP32) {
          float ax[2][8];
#pragma unroll
          for (int rb = 0; rb < 2; ++rb) {
            const float* ap = reinterpret_cast<const float*>(S.A) +
                              (long long)(rbase + 16 * rb + li) * S.lda + kk;
            const f32x4* p = reinterpret_cast<const f32x4*>(ap);
            f32x4 u = p[0], v = p[1];
            ax[rb][0] = u[0]; ax[rb][1] = u[1]; ax[rb][2] = u[2]; ax[rb][3] = u[3];
            ax[rb][4] = v[0]; ax[rb][5] = v[1]; ax[rb][6] = v[2]; ax[rb][7] = v[3];
          }
#pragma unroll
          for (int s = 0; s < 8; ++s)
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
#pragma unroll
              for (int cb = 0; cb < 2; ++cb)
                acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(ax[rb][s], bx[cb][s],
                                                                   acc[rb][cb], 0, 0, 0);
        } else {
          u16x8 bp[2][3];
#pragma unroll
          for (int cb = 0; cb < 2; ++cb) split_pieces<NP>(bx[cb], bp[cb]);
          u16x8 ap[2][3];
#pragma unroll
          for (int rb = 0; rb < 2; ++rb) {
            const unsigned short* a0 = reinterpret_cast<const unsigned short*>(S.A) +
                                       (long long)(rbase + 16 * rb + li) * S.lda + kk;
#pragma unroll
            for (int pc = 0; pc < NP; ++pc)
              ap[rb][pc] = *reinterpret_cast<const u16x8*>(a0 + pc * S.piece);
          }
#pragma unroll
          for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) {
#define SHT_MFMA(ACC, X, Y) \
  ACC = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, X), \
                                                __builtin_bit_cast(bf16x8, Y), ACC, 0, 0, 0)
              if constexpr (MODE == kBF16X6) {
                SHT_MFMA(acc2[rb][cb], ap[rb][2], bp[cb][0]);
                SHT_MFMA(acc2[rb][cb], ap[rb][1], bp[cb][1]);
                SHT_MFMA(acc2[rb][cb], ap[rb][0], bp[cb][2]);
              }
              SHT_MFMA(acc2[rb][cb], ap[rb][1], bp[cb][0]);
              SHT_MFMA(acc2[rb][cb], ap[rb][0], bp[cb][1]);
              SHT_MFMA(acc[rb][cb], ap[rb][0], bp[cb][0]);
#undef SHT_MFMA
            }
        }
      }
    }
    // Store: lane holds rows 4*kg + q (q = 0..3) of column li in each 16x16 tile.
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int cb = 0; cb < 2; ++cb) {
        const int col = cbase + 16 * cb + li;
        if (col >= P.N) continue;
        const int row = rbase + 16 * rb + 4 * kg;
        f32x4 v = acc[rb][cb];
        if constexpr (MODE != kFP32) v = v + acc2[rb][cb];
        float* cp = P.C + (long long)row * P.crs + (long long)col * P.ccs;
        if (P.crs == 1) {
          *reinterpret_cast<f32x4*>(cp) = v;
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) cp[(long long)q * P.crs] = v[q];
        }
      }
  }
}

}  // namespace sht

extern "C" hipError_t sht_launch_legendre_gemm(int mode, const sht::GemmProb* probs,
                                               const int4* tiles, int ntiles, int grid,
                                               hipStream_t s) {
  if (ntiles <= 0) return hipSuccess;
  sht::GemmWork w{probs, tiles, ntiles};
  dim3 g(grid), b(256);
  switch (mode) {
    case sht::kFP32:
      hipLaunchKernelGGL(sht::legendre_gemm_kernel<sht::kFP32>, g, b, 0, s, w);
      break;
    case sht::kBF16X3:
      hipLaunchKernelGGL(sht::legendre_gemm_kernel<sht::kBF16X3>, g, b, 0, s, w);
      break;
    case sht::kBF16X6:
      hipLaunchKernelGGL(sht::legendre_gemm_kernel<sht::kBF16X6>, g, b, 0, s, w);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
