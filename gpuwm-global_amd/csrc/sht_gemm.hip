// Grouped Legendre GEMM over the triangular m-schedule (one persistent launch
// per direction).
//
// Every Legendre contraction of a batch is a list of work items: a panel (a
// group of up to kMT = 4 consecutive orders m of one job and one output
// parity, times one 64-row tile) and a first column (32 columns per item).
// Per order m the group holds
//     C_m[r][c] = sum_seg sum_k A_seg,m[r][k] * B_seg,m[k][c]
// with A a Legendre table block and B a data block; two segments express the
// vector operators (P*m/cos and dP/dphi against differently arranged data) as
// one accumulation.  The host lists each panel's k-steps (one order, 32 k) in
// execution order, with byte addresses and valid extents, so the kernel does
// no schedule arithmetic: panels and steps come in through the scalar cache,
// tiles through bounds-checked buffer loads (reads past an extent return 0,
// which is also how row and column edges are handled).
//
// Pipeline: two register stages and two LDS buffers.  While the wavefronts
// run step j's MFMAs out of LDS, the loads of steps j+1 and j+2 are in
// flight; step j+1 is then written to the other LDS buffer and one barrier
// ends the step.  Each panel's list ends in two null steps (extent 0), so
// every load is issued unconditionally.
//
// A step may read its B columns through a map (reverse each group of four,
// negate some of them): the vector operators' second arrangement of the data
// is such a map of the first, so only one copy is stored.
//
// SYN (synthesis): rows = folded latitudes, k = degrees, C = waist
//   W[c][m/4][p][i][m%4].  Each lane keeps four order accumulators; the tile
//   is stored as 4 latitudes x 4 orders = 64 contiguous bytes per lane.
// ANA (analysis): rows = degrees, k = folded latitudes.  The waist puts the
//   four orders of one latitude side by side, so one B tile (32 latitudes x
//   32 columns x 4 orders) is loaded once per k-chunk and serves the steps
//   of all four orders; steps flagged with a B extent start a new k-chunk.
//
// Block tile 64 x 32, four wavefronts of 32 x 16 (two 16x16 MFMA row blocks).
// Precision modes:
//   FP32   v_mfma_f32_16x16x4_f32, fp32 tables and data.
//   BF16X3 v_mfma_f32_16x16x32_bf16 on hi+lo pieces: a0b0 + a0b1 + a1b0.
//   BF16X6 three pieces each: a0b0 + (a0b1 + a1b0) + (a0b2 + a1b1 + a2b0),
//          every term down to 2^-16 relative; the five small products
//          accumulate apart from a0b0 and are added once at the end.
// Tables are pre-split on the device from FP64 (sht_tables.hip); data are
// split into pieces (round-to-nearest-even) once per block, when a tile is
// written to LDS.  Lane l of a 16x16 MFMA: (l & 15) is the row/column and
// (l >> 4) selects a run of 8 consecutive k; A and B use the same k mapping.
#include <type_traits>

#include "sht_kernels.h"

namespace sht {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) unsigned short u16x8;
typedef __attribute__((ext_vector_type(4))) unsigned short u16x4;
typedef __attribute__((ext_vector_type(2))) unsigned short u16x2;
typedef __attribute__((ext_vector_type(4))) float f32x4;

// Address spaces of the device pass (the host pass only parses the kernels).
#if defined(__HIP_DEVICE_COMPILE__)
#define SHT_CONST __attribute__((address_space(4)))   // scalar-cache loads
#define SHT_GLOBAL __attribute__((address_space(1)))  // global (not flat) stores
#else
#define SHT_CONST
#define SHT_GLOBAL
#endif

constexpr int kBK = 32;            // k per step
constexpr int kBM = 64, kBN = 32;  // block tile
constexpr int kLdF = kBK + 4;      // fp32 LDS row stride (floats)
constexpr int kLdH = kBK + 8;      // bf16 LDS row stride (elements)
constexpr unsigned kRsrcWord3 = 0x00020000u;  // raw buffer, 32-bit data format

struct GemmWork {
  const GemmPanel* panels;
  const GemmStep* steps;
  const int2* tiles;  // {panel, c0}; panel < 0 = idle padding
  int ntiles;
};

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(unsigned long long addr,
                                                            unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void*>(addr), 0, (int)bytes,
                                           (int)kRsrcWord3);
}
__device__ __forceinline__ f32x4 load4(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
}

__device__ __forceinline__ unsigned short f2bf(float f) {
  __bf16 h = (__bf16)f;  // round to nearest even
  return __builtin_bit_cast(unsigned short, h);
}
__device__ __forceinline__ float bf2f(unsigned short b) {
  return __uint_as_float(((unsigned int)b) << 16);
}

// x = p0 + p1 (+ p2) in bf16 pieces, each the nearest bf16 of what is left.
template <int NP>
__device__ __forceinline__ void split1(float x, unsigned short (&p)[3]) {
  p[0] = f2bf(x);
  if constexpr (NP >= 2) {
    const float r = x - bf2f(p[0]);
    p[1] = f2bf(r);
    if constexpr (NP >= 3) p[2] = f2bf(r - bf2f(p[1]));
  }
}

#define SHT_MFMA_BF(ACC, X, Y)                                                                 \
  ACC = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, X),                \
                                                __builtin_bit_cast(bf16x8, Y), ACC, 0, 0, 0)

// acc (+acc2) += A-fragment x B-fragment for one 16x16 tile.
template <int MODE>
__device__ __forceinline__ void mfma_tile(f32x4& acc, f32x4& acc2, const float (&ax)[8],
                                          const u16x8 (&ap)[3], const float (&bx)[8],
                                          const u16x8 (&bp)[3]) {
  if constexpr (MODE == kFP32) {
#pragma unroll
    for (int s = 0; s < 8; ++s) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(ax[s], bx[s], acc, 0, 0, 0);
  } else {
    if constexpr (MODE == kBF16X6) {
      SHT_MFMA_BF(acc2, ap[2], bp[0]);
      SHT_MFMA_BF(acc2, ap[1], bp[1]);
      SHT_MFMA_BF(acc2, ap[0], bp[2]);
    }
    SHT_MFMA_BF(acc2, ap[1], bp[0]);
    SHT_MFMA_BF(acc2, ap[0], bp[1]);
    SHT_MFMA_BF(acc, ap[0], bp[0]);
  }
}

template <int MODE, int DIR>
struct GemmTraits {
  static constexpr bool kF32 = MODE == kFP32;
  static constexpr int NP = MODE == kBF16X6 ? 3 : 2;  // pieces (bf16 modes)
  static constexpr int A_LOADS = kF32 ? 2 : NP;       // 16-byte loads per thread per A tile
  static constexpr int B_LOADS = DIR == 0 ? 1 : 4;    // per B tile (ANA: four orders)
  static constexpr int BZ = DIR == 0 ? 1 : kMT;       // orders per B tile
  // LDS, in elements of the storage type
  static constexpr int A_F = kBM * kLdF;              // fp32 A tile
  static constexpr int A_H = NP * kBM * kLdH;         // bf16 A tile, NP pieces
  static constexpr int B_F = BZ * kBN * kLdF;
  static constexpr int B_H = NP * BZ * kBN * kLdH;
  static constexpr int B_BUFS = DIR == 0 ? 2 : 1;     // ANA keeps one k-chunk resident
};

template <int MODE, int DIR>
__global__ void __launch_bounds__(256) legendre_gemm_kernel(GemmWork w) {
  using TR = GemmTraits<MODE, DIR>;
  constexpr bool F32 = TR::kF32;
  constexpr int NP = TR::NP;
  __shared__ __attribute__((aligned(16))) float sAf[2][F32 ? TR::A_F : 4];
  __shared__ __attribute__((aligned(16))) unsigned short sAh[2][F32 ? 8 : TR::A_H];
  __shared__ __attribute__((aligned(16))) float sBf[TR::B_BUFS][F32 ? TR::B_F : 4];
  __shared__ __attribute__((aligned(16))) unsigned short sBh[TR::B_BUFS][F32 ? 8 : TR::B_H];

  const SHT_CONST GemmPanel* panels = (const SHT_CONST GemmPanel*)w.panels;
  const SHT_CONST GemmStep* steps = (const SHT_CONST GemmStep*)w.steps;
  const SHT_CONST int2* tiles = (const SHT_CONST int2*)w.tiles;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, kg = lane >> 4;
  const int wr = wave >> 1, wc = wave & 1;

  for (int t = blockIdx.x; t < w.ntiles; t += gridDim.x) {
    const int2 tile = tiles[t];
    if (tile.x < 0) continue;
    const GemmPanel P = panels[tile.x];
    const int c0 = tile.y;
    const int ncol = min(kBN, P.N - c0);
    const unsigned long long coloff = (unsigned long long)c0 * P.bcs * 4;
    const int step0 = P.step0, nsteps = P.nsteps;

    // ---- stage registers and their loads ----
    f32x4 ra[2][TR::A_LOADS], rb[2][TR::B_LOADS];
    auto issue = [&](f32x4(&A)[TR::A_LOADS], f32x4(&B)[TR::B_LOADS], int j) {
      const GemmStep s = steps[step0 + j];
      if constexpr (F32) {
        const __amdgpu_buffer_rsrc_t r = make_rsrc(s.a, s.a_bytes);
#pragma unroll
        for (int u = 0; u < 2; ++u) A[u] = load4(r, (unsigned)(tid + 256 * u) * 16u);
      } else {
#pragma unroll
        for (int pc = 0; pc < NP; ++pc)
          A[pc] = load4(make_rsrc(s.a + (unsigned long long)(pc * P.piece_bytes), s.a_bytes),
                        (unsigned)tid * 16u);
      }
      if constexpr (DIR == 0) {
        const unsigned bytes =
            s.b_klen > 0 ? (unsigned)(((ncol - 1) * P.bcs + s.b_klen) * 4) : 0u;
        const __amdgpu_buffer_rsrc_t r = make_rsrc(s.b + coloff, bytes);
        const int src = (tid >> 3) ^ ((s.cmap & 1) ? 3 : 0);
        B[0] = load4(r, (unsigned)(src * P.bcs + (tid & 7) * 4) * 4u);
      } else {
        const unsigned bytes =
            s.b_klen > 0 ? (unsigned)(((ncol - 1) * P.bcs + s.b_klen * kMT) * 4) : 0u;
        const __amdgpu_buffer_rsrc_t r = make_rsrc(s.b + coloff, bytes);
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int q = tid + 256 * u;  // column q >> 4, latitude pair q & 15
          const int src = (q >> 4) ^ ((s.cmap & 1) ? 3 : 0);
          const unsigned off = (unsigned)(src * P.bcs + (q & 15) * 2 * kMT) * 4u;
          B[2 * u] = load4(r, off);
          B[2 * u + 1] = load4(r, off + 16u);
        }
      }
    };
    // A stage -> LDS buffer
    auto put_a = [&](const f32x4(&A)[TR::A_LOADS], int buf) {
      if constexpr (F32) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int q = tid + 256 * u;
          *reinterpret_cast<f32x4*>(&sAf[buf][(q >> 3) * kLdF + (q & 7) * 4]) = A[u];
        }
      } else {
#pragma unroll
        for (int pc = 0; pc < NP; ++pc)
          *reinterpret_cast<f32x4*>(&sAh[buf][(pc * kBM + (tid >> 2)) * kLdH + (tid & 3) * 8]) =
              A[pc];
      }
    };
    // B stage -> LDS (split into pieces for the bf16 modes); `cmap` bits
    // 1..4 negate columns 0..3 of each group of four.
    auto put_b = [&](const f32x4(&B)[TR::B_LOADS], int buf, int cmap) {
      if constexpr (DIR == 0) {
        const int col = tid >> 3, kq = (tid & 7) * 4;
        const f32x4 v = ((cmap >> (1 + (col & 3))) & 1) ? -B[0] : B[0];
        if constexpr (F32) {
          *reinterpret_cast<f32x4*>(&sBf[buf][col * kLdF + kq]) = v;
        } else {
          u16x4 pv[3];
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            unsigned short p[3];
            split1<NP>(v[s], p);
#pragma unroll
            for (int pc = 0; pc < NP; ++pc) pv[pc][s] = p[pc];
          }
#pragma unroll
          for (int pc = 0; pc < NP; ++pc)
            *reinterpret_cast<u16x4*>(&sBh[buf][(pc * kBN + col) * kLdH + kq]) = pv[pc];
        }
      } else {
        (void)buf;
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int q = tid + 256 * u;
          const int col = q >> 4, k2 = (q & 15) * 2;
          const bool neg = (cmap >> (1 + (col & 3))) & 1;
          const f32x4 v0 = neg ? -B[2 * u] : B[2 * u], v1 = neg ? -B[2 * u + 1] : B[2 * u + 1];
#pragma unroll
          for (int z = 0; z < kMT; ++z) {
            if constexpr (F32) {
              *reinterpret_cast<float2*>(&sBf[0][(z * kBN + col) * kLdF + k2]) =
                  float2{v0[z], v1[z]};
            } else {
              unsigned short p0[3], p1[3];
              split1<NP>(v0[z], p0);
              split1<NP>(v1[z], p1);
#pragma unroll
              for (int pc = 0; pc < NP; ++pc)
                *reinterpret_cast<u16x2*>(&sBh[0][((pc * kMT + z) * kBN + col) * kLdH + k2]) =
                    u16x2{p0[pc], p1[pc]};
            }
          }
        }
      }
    };

    f32x4 acc[kMT][2], acc2[kMT][2];
#pragma unroll
    for (int a = 0; a < kMT; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
        acc2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
      }

    // one step out of LDS buffer `buf` into the accumulators of order z
    auto compute = [&](int buf, int z) {
      const int bbuf = DIR == 0 ? buf : 0;
      const int bz = DIR == 0 ? 0 : z;
      float ax[2][8], bx[8];
      u16x8 ap[2][3], bp[3];
#pragma unroll
      for (int r2 = 0; r2 < 2; ++r2) {
        const int row = wr * 32 + r2 * 16 + li;
        if constexpr (F32) {
          const f32x4* p = reinterpret_cast<const f32x4*>(&sAf[buf][row * kLdF + 8 * kg]);
          const f32x4 u = p[0], v = p[1];
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            ax[r2][s] = u[s];
            ax[r2][4 + s] = v[s];
          }
        } else {
#pragma unroll
          for (int pc = 0; pc < NP; ++pc)
            ap[r2][pc] =
                *reinterpret_cast<const u16x8*>(&sAh[buf][(pc * kBM + row) * kLdH + 8 * kg]);
        }
      }
      const int col = wc * 16 + li;
      if constexpr (F32) {
        const f32x4* p =
            reinterpret_cast<const f32x4*>(&sBf[bbuf][(bz * kBN + col) * kLdF + 8 * kg]);
        const f32x4 u = p[0], v = p[1];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          bx[s] = u[s];
          bx[4 + s] = v[s];
        }
      } else {
#pragma unroll
        for (int pc = 0; pc < NP; ++pc)
          bp[pc] = *reinterpret_cast<const u16x8*>(
              &sBh[bbuf][((pc * TR::BZ + bz) * kBN + col) * kLdH + 8 * kg]);
      }
#pragma unroll
      for (int zz = 0; zz < kMT; ++zz) {
        if (zz != z) continue;
#pragma unroll
        for (int r2 = 0; r2 < 2; ++r2)
          mfma_tile<MODE>(acc[zz][r2], acc2[zz][r2], ax[r2], ap[r2], bx, bp);
      }
    };

    // ---- prologue: steps 0 and 1 in flight, step 0 in LDS buffer 0 ----
    issue(ra[0], rb[0], 0);
    issue(ra[1], rb[1], 1);
    put_a(ra[0], 0);
    if (DIR == 0 || steps[step0].b_klen > 0) put_b(rb[0], 0, steps[step0].cmap);
    __syncthreads();

    // ---- main loop, two steps per trip (register stage = step parity) ----
    auto body = [&](auto S, int j) {
      constexpr int s = decltype(S)::value;
      issue(ra[s], rb[s], j + 2);
      compute(s, steps[step0 + j].z);
      if (j + 1 < nsteps) {
        const GemmStep nx = steps[step0 + j + 1];
        if constexpr (DIR == 0) {
          put_b(rb[s ^ 1], s ^ 1, nx.cmap);
        } else {
          if (nx.b_klen > 0) {
            __syncthreads();  // every wavefront is done with the resident B tile
            put_b(rb[s ^ 1], 0, nx.cmap);
          }
        }
        put_a(ra[s ^ 1], s ^ 1);
      }
      __syncthreads();
    };
    for (int j = 0; j < nsteps; j += 2) {
      body(std::integral_constant<int, 0>{}, j);
      if (j + 1 < nsteps) body(std::integral_constant<int, 1>{}, j + 1);
    }

    // ---- epilogue ----
#pragma unroll
    for (int r2 = 0; r2 < 2; ++r2) {
      const int col = wc * 16 + li;
      const int row = wr * 32 + r2 * 16 + 4 * kg;
      if (col >= ncol) continue;
      if constexpr (DIR == 0) {
        // rows row..row+3 x orders 0..3 = 16 contiguous floats of the waist
        if (row >= P.rows[0]) continue;
        SHT_GLOBAL float* cp = reinterpret_cast<SHT_GLOBAL float*>(P.c[0]) +
                               (long long)(c0 + col) * P.ccs + (long long)row * kMT;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          f32x4 o;
#pragma unroll
          for (int zz = 0; zz < kMT; ++zz) {
            f32x4 v = acc[zz][r2];
            if constexpr (!F32) v = v + acc2[zz][r2];
            o[zz] = v[q];
          }
          *reinterpret_cast<SHT_GLOBAL f32x4*>(cp + q * kMT) = o;
        }
      } else {
#pragma unroll
        for (int zz = 0; zz < kMT; ++zz) {
          if (row >= P.rows[zz]) continue;
          f32x4 v = acc[zz][r2];
          if constexpr (!F32) v = v + acc2[zz][r2];
          *reinterpret_cast<SHT_GLOBAL f32x4*>(reinterpret_cast<SHT_GLOBAL float*>(P.c[zz]) +
                                               (long long)(c0 + col) * P.ccs + row) = v;
        }
      }
    }
  }
}

#undef SHT_MFMA_BF

}  // namespace sht

extern "C" hipError_t sht_launch_legendre_gemm(int mode, int dir, const sht::GemmPanel* panels,
                                               const sht::GemmStep* steps, const int2* tiles,
                                               int ntiles, int grid, hipStream_t s) {
  if (ntiles <= 0) return hipSuccess;
  sht::GemmWork w{panels, steps, tiles, ntiles};
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

extern "C" int sht_gemm_tile_rows(int dir) { (void)dir; return sht::kBM; }
extern "C" int sht_gemm_tile_cols(int dir) { (void)dir; return sht::kBN; }
