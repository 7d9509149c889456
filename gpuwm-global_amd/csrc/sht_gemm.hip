// Grouped Legendre GEMM over the triangular m-schedule (one persistent launch).
//
// Every Legendre contraction of a batch is an entry of a work list: a group
// of up to eight consecutive orders m (one m-tile of the waist layout) for
// one job and one output parity, times one row/column tile.  Each workgroup
// walks the list with a grid stride, so one launch covers all fields, levels,
// members and truncations of the batch.  Per order m the group holds
//     C_m[r][c] = sum_seg sum_k A_seg,m[r][k] * B_seg,m[k][c]
// with A a Legendre table block (k contiguous, rows and k zero padded to 32)
// and B a data block.  Two segments express the vector operators (P*m/cos
// and dP/dphi against differently arranged data) as one accumulation.
//
// Direction SYN: B = packed coefficients (k = degree, contiguous per order),
//   C = waist W[c][m/8][p][i][m%8]; the eight orders' accumulators are kept
//   in registers, so each lane stores 4 latitudes x 8 orders = 128
//   contiguous bytes per column.
// Direction ANA: B = waist (k = latitude); each lane loads 8 latitudes x 8
//   orders = 256 contiguous bytes per column and picks each order's fragment
//   in registers; C = packed coefficients (4 contiguous degrees per lane).
//
// Precision modes:
//   FP32   v_mfma_f32_16x16x4_f32, fp32 tables and data.
//   BF16X3 v_mfma_f32_16x16x32_bf16 on hi+lo pieces: a0b0 + a0b1 + a1b0.
//   BF16X6 three pieces each: a0b0 + (a0b1 + a1b0) + (a0b2 + a1b1 + a2b0),
//          every term down to 2^-16 relative; the five small products
//          accumulate apart from a0b0 and are added once at the end.
// Tables are pre-split on the device from FP64 (sht_tables.hip); data are
// split here, in registers, by round-to-nearest-even conversion.
//
// Tile: 256 threads = 4 wavefronts in a 2 x 2 arrangement, each wavefront
// RT x CT MFMA 16x16 tiles.  Lane l: (l & 15) is the row/column, (l >> 4)
// selects a run of 8 consecutive k; A and B fragments use the same k
// mapping, so each MFMA sums matching k.
#include "sht_kernels.h"

namespace sht {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) unsigned short u16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

constexpr int kRT = 2;  // MFMA row tiles per wavefront
constexpr int kCT = 1;  // MFMA column tiles per wavefront
constexpr int kWR = 2;  // wavefront rows per block
constexpr int kWC = 2;  // wavefront columns per block

struct GemmWork {
  const GemmGroup* groups;
  const int4* tiles;  // {group, r0, c0, unused}
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

#define SHT_MFMA_BF(ACC, X, Y)                                                                 \
  ACC = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, X),                \
                                                __builtin_bit_cast(bf16x8, Y), ACC, 0, 0, 0)

template <int MODE, int DIR>
__global__ void __launch_bounds__(256) legendre_gemm_kernel(GemmWork w) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int li = lane & 15;
  const int kg = lane >> 4;
  constexpr int NP = MODE == kBF16X6 ? 3 : 2;
  constexpr bool SPLIT = MODE != kFP32;

  for (int t = blockIdx.x; t < w.ntiles; t += gridDim.x) {
    const int4 tile = w.tiles[t];
    const GemmGroup& G = w.groups[tile.x];
    const int rbase = tile.y + 16 * kRT * (wave / kWC);
    const int cbase = tile.z + 16 * kCT * (wave % kWC);
    if (rbase >= G.Rmax || cbase >= G.N) continue;  // wave-uniform

    f32x4 acc[kMT][kRT][kCT];
    f32x4 acc2[kMT][kRT][kCT];
#pragma unroll
    for (int mm = 0; mm < kMT; ++mm)
#pragma unroll
      for (int rb = 0; rb < kRT; ++rb)
#pragma unroll
        for (int cb = 0; cb < kCT; ++cb) {
          acc[mm][rb][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
          acc2[mm][rb][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
        }

    for (int sg = 0; sg < G.nseg; ++sg) {
      const int bks = G.bks[sg];
      const int bcs = G.bcs[sg];
      for (int k0 = 0; k0 < G.Kmax[sg]; k0 += 32) {
        const int kk = k0 + 8 * kg;
        // ANA: one 256-byte load per column covers all eight orders.
        float wb[kCT][64];
        if constexpr (DIR == 1) {
#pragma unroll
          for (int cb = 0; cb < kCT; ++cb) {
            const int col = cbase + 16 * cb + li;
            if (col < G.N) {
              const f32x4* p = reinterpret_cast<const f32x4*>(
                  G.B[sg][0] + (long long)col * bcs + (long long)kk * kMT);
#pragma unroll
              for (int v = 0; v < 16; ++v) {
                f32x4 x = p[v];
                wb[cb][4 * v] = x[0];
                wb[cb][4 * v + 1] = x[1];
                wb[cb][4 * v + 2] = x[2];
                wb[cb][4 * v + 3] = x[3];
              }
            } else {
#pragma unroll
              for (int v = 0; v < 64; ++v) wb[cb][v] = 0.f;
            }
          }
        }
#pragma unroll
        for (int mm = 0; mm < kMT; ++mm) {
          if (mm >= G.mcount || k0 >= G.K[sg][mm] || rbase >= G.R[mm]) continue;
          // B fragments for this order
          float bx[kCT][8];
#pragma unroll
          for (int cb = 0; cb < kCT; ++cb) {
            if constexpr (DIR == 1) {
#pragma unroll
              for (int s = 0; s < 8; ++s) bx[cb][s] = wb[cb][s * kMT + mm];
            } else {
              const int col = cbase + 16 * cb + li;
              if (col < G.N) {
                const f32x4* p = reinterpret_cast<const f32x4*>(
                    G.B[sg][mm] + (long long)col * bcs + (long long)kk * bks);
                f32x4 u = p[0], v = p[1];
                bx[cb][0] = u[0]; bx[cb][1] = u[1]; bx[cb][2] = u[2]; bx[cb][3] = u[3];
                bx[cb][4] = v[0]; bx[cb][5] = v[1]; bx[cb][6] = v[2]; bx[cb][7] = v[3];
              } else {
#pragma unroll
                for (int s = 0; s < 8; ++s) bx[cb][s] = 0.f;
              }
            }
          }
          const int lda = G.lda[sg][mm];
          if constexpr (!SPLIT) {
            float ax[kRT][8];
#pragma unroll
            for (int rb = 0; rb < kRT; ++rb) {
              const f32x4* p = reinterpret_cast<const f32x4*>(
                  reinterpret_cast<const float*>(G.A[sg][mm]) +
                  (long long)(rbase + 16 * rb + li) * lda + kk);
              f32x4 u = p[0], v = p[1];
              ax[rb][0] = u[0]; ax[rb][1] = u[1]; ax[rb][2] = u[2]; ax[rb][3] = u[3];
              ax[rb][4] = v[0]; ax[rb][5] = v[1]; ax[rb][6] = v[2]; ax[rb][7] = v[3];
            }
#pragma unroll
            for (int s = 0; s < 8; ++s)
#pragma unroll
              for (int rb = 0; rb < kRT; ++rb)
#pragma unroll
                for (int cb = 0; cb < kCT; ++cb)
                  acc[mm][rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                      ax[rb][s], bx[cb][s], acc[mm][rb][cb], 0, 0, 0);
          } else {
            u16x8 bp[kCT][3];
#pragma unroll
            for (int cb = 0; cb < kCT; ++cb) split_pieces<NP>(bx[cb], bp[cb]);
            u16x8 ap[kRT][3];
#pragma unroll
            for (int rb = 0; rb < kRT; ++rb) {
              const unsigned short* a0 = reinterpret_cast<const unsigned short*>(G.A[sg][mm]) +
                                         (long long)(rbase + 16 * rb + li) * lda + kk;
#pragma unroll
              for (int pc = 0; pc < NP; ++pc)
                ap[rb][pc] = *reinterpret_cast<const u16x8*>(a0 + pc * G.piece);
            }
#pragma unroll
            for (int rb = 0; rb < kRT; ++rb)
#pragma unroll
              for (int cb = 0; cb < kCT; ++cb) {
                if constexpr (MODE == kBF16X6) {
                  SHT_MFMA_BF(acc2[mm][rb][cb], ap[rb][2], bp[cb][0]);
                  SHT_MFMA_BF(acc2[mm][rb][cb], ap[rb][1], bp[cb][1]);
                  SHT_MFMA_BF(acc2[mm][rb][cb], ap[rb][0], bp[cb][2]);
                }
                SHT_MFMA_BF(acc2[mm][rb][cb], ap[rb][1], bp[cb][0]);
                SHT_MFMA_BF(acc2[mm][rb][cb], ap[rb][0], bp[cb][1]);
                SHT_MFMA_BF(acc[mm][rb][cb], ap[rb][0], bp[cb][0]);
              }
          }
        }
      }
    }

    // Epilogue.  Lane holds rows 4*kg + q (q = 0..3) of column li per 16x16 tile.
#pragma unroll
    for (int rb = 0; rb < kRT; ++rb)
#pragma unroll
      for (int cb = 0; cb < kCT; ++cb) {
        const int col = cbase + 16 * cb + li;
        if (col >= G.N) continue;
        const int row = rbase + 16 * rb + 4 * kg;
        if constexpr (DIR == 0) {
          // waist: rows row..row+3 x orders 0..7 are 32 contiguous floats
          float* cp = G.C[0] + (long long)col * G.ccs + (long long)row * kMT;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            float o[kMT];
#pragma unroll
            for (int mm = 0; mm < kMT; ++mm) {
              f32x4 v = acc[mm][rb][cb];
              if constexpr (SPLIT) v = v + acc2[mm][rb][cb];
              o[mm] = (mm < G.mcount) ? v[q] : 0.f;
            }
            f32x4* d = reinterpret_cast<f32x4*>(cp + q * kMT);
            d[0] = f32x4{o[0], o[1], o[2], o[3]};
            d[1] = f32x4{o[4], o[5], o[6], o[7]};
          }
        } else {
#pragma unroll
          for (int mm = 0; mm < kMT; ++mm) {
            if (mm >= G.mcount || row >= G.R[mm]) continue;
            f32x4 v = acc[mm][rb][cb];
            if constexpr (SPLIT) v = v + acc2[mm][rb][cb];
            *reinterpret_cast<f32x4*>(G.C[mm] + (long long)col * G.ccs + row) = v;
          }
        }
      }
  }
}

#undef SHT_MFMA_BF

}  // namespace sht

extern "C" hipError_t sht_launch_legendre_gemm(int mode, int dir, const sht::GemmGroup* groups,
                                               const int4* tiles, int ntiles, int grid,
                                               hipStream_t s) {
  if (ntiles <= 0) return hipSuccess;
  sht::GemmWork w{groups, tiles, ntiles};
  dim3 g(grid), b(256);
#define SHT_LAUNCH(M, D) hipLaunchKernelGGL((sht::legendre_gemm_kernel<M, D>), g, b, 0, s, w)
  if (dir == 0) {
    if (mode == sht::kFP32) SHT_LAUNCH(sht::kFP32, 0);
    else if (mode == sht::kBF16X3) SHT_LAUNCH(sht::kBF16X3, 0);
    else if (mode == sht::kBF16X6) SHT_LAUNCH(sht::kBF16X6, 0);
    else return hipErrorInvalidValue;
  } else {
    if (mode == sht::kFP32) SHT_LAUNCH(sht::kFP32, 1);
    else if (mode == sht::kBF16X3) SHT_LAUNCH(sht::kBF16X3, 1);
    else if (mode == sht::kBF16X6) SHT_LAUNCH(sht::kBF16X6, 1);
    else return hipErrorInvalidValue;
  }
#undef SHT_LAUNCH
  return hipGetLastError();
}

extern "C" int sht_gemm_tile_rows() { return 16 * sht::kRT * sht::kWR; }
extern "C" int sht_gemm_tile_cols() { return 16 * sht::kCT * sht::kWC; }
