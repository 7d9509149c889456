// Grouped Legendre GEMM over the triangular m-schedule (one persistent launch
// per direction).
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
// Both directions stream 32-deep k-steps through double-buffered LDS: all
// 256 threads fetch the next step's A and B tiles into registers while the
// four wavefronts run the current step's MFMAs, then store them and meet at
// one barrier per step.
//
// SYN (synthesis): rows = folded latitudes, k = degrees (packed, contiguous
//   per order), C = waist W[c][m/8][p][i][m%8].  The loop runs k outer,
//   order inner, with eight accumulator sets, so the finished tile is stored
//   as 4 latitudes x 8 orders = 128 contiguous bytes per lane and column
//   (storing each order separately writes every waist line eight times).
//   Block tile 64 x 32, each wavefront 32 x 16.
// ANA (analysis): rows = degrees, k = folded latitudes from the waist, whose
//   m-tile layout puts eight orders' values for one latitude side by side:
//   one 32-latitude B tile (1 KB per column) serves all eight orders, so the
//   loop runs k outer, order inner, with eight accumulator sets.  Block tile
//   64 x 32, each wavefront 32 x 16.
//
// Precision modes:
//   FP32   v_mfma_f32_16x16x4_f32, fp32 tables and data.
//   BF16X3 v_mfma_f32_16x16x32_bf16 on hi+lo pieces: a0b0 + a0b1 + a1b0.
//   BF16X6 three pieces each: a0b0 + (a0b1 + a1b0) + (a0b2 + a1b1 + a2b0),
//          every term down to 2^-16 relative; the five small products
//          accumulate apart from a0b0 and are added once at the end.
// Tables are pre-split on the device from FP64 (sht_tables.hip); data are
// split when their fragments are read from LDS, by round-to-nearest-even
// conversion.  Lane l of a 16x16 MFMA: (l & 15) is the row/column, (l >> 4)
// selects a run of 8 consecutive k; A and B fragments use the same k mapping.
#include "sht_kernels.h"

namespace sht {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) unsigned short u16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

constexpr int kBK = 32;       // k per step
constexpr int kPadF = 4;      // LDS row padding, floats
constexpr int kPadH = 8;      // LDS row padding, bf16
constexpr int kLdF = kBK + kPadF;  // fp32 LDS row stride
constexpr int kLdH = kBK + kPadH;  // bf16 LDS row stride

constexpr int kSynBM = 64, kSynBN = 32;
constexpr int kAnaBM = 64, kAnaBN = 32;

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

// A tile staging: BM rows x 32 k.  FP32: one 8-float run per thread and row
// (BM*4 runs); bf16: one 8-element run per thread, row and piece.
template <int MODE, int BM>
struct ATile {
  static constexpr int NP = MODE == kBF16X6 ? 3 : (MODE == kBF16X3 ? 2 : 1);
  static constexpr int RUNS = BM * 4;                 // 8-element runs per piece
  static constexpr int PER_THREAD = RUNS / 256;       // runs per thread per piece
  static constexpr int LDS_ELEMS = MODE == kFP32 ? BM * kLdF : NP * BM * kLdH;
  f32x4 f[PER_THREAD][2];
  u16x8 h[PER_THREAD][3];

  __device__ __forceinline__ void load(const void* A, int lda, long long piece, int r0, int R,
                                       int k0) {
#pragma unroll
    for (int u = 0; u < PER_THREAD; ++u) {
      const int run = threadIdx.x + 256 * u;
      const int row = run >> 2;
      const int kq = (run & 3) * 8;
      const bool ok = r0 + row < R;
      if constexpr (MODE == kFP32) {
        f[u][0] = f32x4{0.f, 0.f, 0.f, 0.f};
        f[u][1] = f32x4{0.f, 0.f, 0.f, 0.f};
        if (ok) {
          const f32x4* p = reinterpret_cast<const f32x4*>(
              reinterpret_cast<const float*>(A) + (long long)(r0 + row) * lda + k0 + kq);
          f[u][0] = p[0];
          f[u][1] = p[1];
        }
      } else {
#pragma unroll
        for (int pc = 0; pc < NP; ++pc) h[u][pc] = u16x8{};
        if (ok) {
          const unsigned short* a0 = reinterpret_cast<const unsigned short*>(A) +
                                     (long long)(r0 + row) * lda + k0 + kq;
#pragma unroll
          for (int pc = 0; pc < NP; ++pc) h[u][pc] = *reinterpret_cast<const u16x8*>(a0 + pc * piece);
        }
      }
    }
  }
  __device__ __forceinline__ void store(float* sf, unsigned short* sh) const {
#pragma unroll
    for (int u = 0; u < PER_THREAD; ++u) {
      const int run = threadIdx.x + 256 * u;
      const int row = run >> 2;
      const int kq = (run & 3) * 8;
      if constexpr (MODE == kFP32) {
        f32x4* d = reinterpret_cast<f32x4*>(sf + row * kLdF + kq);
        d[0] = f[u][0];
        d[1] = f[u][1];
      } else {
#pragma unroll
        for (int pc = 0; pc < NP; ++pc)
          *reinterpret_cast<u16x8*>(sh + (pc * BM + row) * kLdH + kq) = h[u][pc];
      }
    }
  }
};

// Fragment reads from LDS.
template <int MODE, int BM>
__device__ __forceinline__ void read_a(const float* sf, const unsigned short* sh, int row, int kg,
                                       float (&ax)[8], u16x8 (&ap)[3]) {
  constexpr int NP = MODE == kBF16X6 ? 3 : 2;
  if constexpr (MODE == kFP32) {
    const f32x4* p = reinterpret_cast<const f32x4*>(sf + row * kLdF + 8 * kg);
    f32x4 u = p[0], v = p[1];
    ax[0] = u[0]; ax[1] = u[1]; ax[2] = u[2]; ax[3] = u[3];
    ax[4] = v[0]; ax[5] = v[1]; ax[6] = v[2]; ax[7] = v[3];
  } else {
#pragma unroll
    for (int pc = 0; pc < NP; ++pc)
      ap[pc] = *reinterpret_cast<const u16x8*>(sh + (pc * BM + row) * kLdH + 8 * kg);
  }
}

__device__ __forceinline__ void read_b(const float* sb, int col, int kg, float (&bx)[8]) {
  const f32x4* p = reinterpret_cast<const f32x4*>(sb + col * kLdF + 8 * kg);
  f32x4 u = p[0], v = p[1];
  bx[0] = u[0]; bx[1] = u[1]; bx[2] = u[2]; bx[3] = u[3];
  bx[4] = v[0]; bx[5] = v[1]; bx[6] = v[2]; bx[7] = v[3];
}

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

// ---------------------------------------------------------------------------
// Synthesis kernel
// ---------------------------------------------------------------------------
// k outer, order inner, eight accumulator sets: when the tile is done each
// lane holds 4 latitudes x 8 orders of one column, which the waist layout
// stores as 128 contiguous bytes.  Orders whose packed block is shorter than
// the current k skip the step.
template <int MODE>
__global__ void __launch_bounds__(256) legendre_syn_kernel(GemmWork w) {
  constexpr int BM = kSynBM, BN = kSynBN;
  constexpr int NP = MODE == kBF16X6 ? 3 : 2;
  using AT = ATile<MODE, BM>;
  __shared__ __attribute__((aligned(16))) float sAf[2][MODE == kFP32 ? AT::LDS_ELEMS : 4];
  __shared__ __attribute__((aligned(16))) unsigned short sAh[2][MODE == kFP32 ? 8 : AT::LDS_ELEMS];
  __shared__ __attribute__((aligned(16))) float sB[2][BN * kLdF];

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, kg = lane >> 4;
  const int wr = wave >> 1, wc = wave & 1;

  for (int t = blockIdx.x; t < w.ntiles; t += gridDim.x) {
    const int4 tile = w.tiles[t];
    if (tile.x < 0) continue;  // idle padding of the XCD-interleaved list
    const GemmGroup& G = w.groups[tile.x];
    const int r0 = tile.y, c0 = tile.z;
    // B staging: 32 cols x 32 k = 256 runs of 4 floats, one per thread
    f32x4 bv;
    auto load_b = [&](int s_, int m_, int k_) {
      const int col = threadIdx.x >> 3;
      const int kq = (threadIdx.x & 7) * 4;
      bv = f32x4{0.f, 0.f, 0.f, 0.f};
      if (c0 + col < G.N)
        bv = *reinterpret_cast<const f32x4*>(G.B[s_][m_] + (long long)(c0 + col) * G.bcs[s_] +
                                             k_ + kq);
    };
    auto store_b = [&](float* sb) {
      const int col = threadIdx.x >> 3;
      const int kq = (threadIdx.x & 7) * 4;
      *reinterpret_cast<f32x4*>(sb + col * kLdF + kq) = bv;
    };
    // step cursor (sg, k0, z): the orders z with K[sg][z] > k0, k0 < Kmax[sg]
    auto first_z = [&](int s_, int k_, int z_) {
      while (z_ < G.mcount && k_ >= G.K[s_][z_]) ++z_;
      return z_;
    };
    auto advance = [&](int& s_, int& k_, int& z_) {
      z_ = first_z(s_, k_, z_ + 1);
      while (z_ >= G.mcount) {
        k_ += kBK;
        if (k_ >= G.Kmax[s_]) {
          k_ = 0;
          if (++s_ >= G.nseg) return false;
        }
        z_ = first_z(s_, k_, 0);
      }
      return true;
    };
    int sg = 0, k0 = 0, z = first_z(0, 0, 0);
    AT at;

    f32x4 acc[kMT][2], acc2[kMT][2];
#pragma unroll
    for (int a = 0; a < kMT; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
        acc2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
      }

    // prologue
    at.load(G.A[sg][z], G.lda[sg][z], G.piece, r0, G.R[z], k0);
    load_b(sg, z, k0);
    at.store(sAf[0], sAh[0]);
    store_b(sB[0]);
    __syncthreads();

    int buf = 0;
    bool more = true;
    while (more) {
      int ns = sg, nk = k0, nz = z;
      const bool has_next = advance(ns, nk, nz);
      if (has_next) {
        at.load(G.A[ns][nz], G.lda[ns][nz], G.piece, r0, G.R[nz], nk);
        load_b(ns, nz, nk);
      }
      {
        const float* sa = sAf[buf];
        const unsigned short* sh = sAh[buf];
        float ax[2][8], bx[8];
        u16x8 ap[2][3], bp[3];
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) read_a<MODE, BM>(sa, sh, wr * 32 + rb * 16 + li, kg, ax[rb], ap[rb]);
        read_b(sB[buf], wc * 16 + li, kg, bx);
        if constexpr (MODE != kFP32) split_pieces<NP>(bx, bp);
#pragma unroll
        for (int zz = 0; zz < kMT; ++zz) {
          if (zz != z) continue;
#pragma unroll
          for (int rb = 0; rb < 2; ++rb) mfma_tile<MODE>(acc[zz][rb], acc2[zz][rb], ax[rb], ap[rb], bx, bp);
        }
      }
      if (has_next) {
        at.store(sAf[buf ^ 1], sAh[buf ^ 1]);
        store_b(sB[buf ^ 1]);
      }
      __syncthreads();
      buf ^= 1;
      sg = ns;
      k0 = nk;
      z = nz;
      more = has_next;
    }
    // Store: rows row..row+3 x orders 0..7 are 32 contiguous floats of the
    // waist (orders beyond T+1 in the last tile are written as zeros).
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
      const int col = c0 + wc * 16 + li;
      const int row = r0 + wr * 32 + rb * 16 + 4 * kg;
      if (col >= G.N || row >= G.Rmax) continue;
      float* cp = G.C[0] + (long long)col * G.ccs + (long long)row * kMT;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        float o[kMT];
#pragma unroll
        for (int zz = 0; zz < kMT; ++zz) {
          f32x4 v = acc[zz][rb];
          if constexpr (MODE != kFP32) v = v + acc2[zz][rb];
          o[zz] = v[q];
        }
        f32x4* d = reinterpret_cast<f32x4*>(cp + q * kMT);
        d[0] = f32x4{o[0], o[1], o[2], o[3]};
        d[1] = f32x4{o[4], o[5], o[6], o[7]};
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Analysis kernel
// ---------------------------------------------------------------------------
template <int MODE>
__global__ void __launch_bounds__(256) legendre_ana_kernel(GemmWork w) {
  constexpr int BM = kAnaBM, BN = kAnaBN;
  constexpr int NP = MODE == kBF16X6 ? 3 : 2;
  using AT = ATile<MODE, BM>;
  __shared__ __attribute__((aligned(16))) float sAf[2][MODE == kFP32 ? AT::LDS_ELEMS : 4];
  __shared__ __attribute__((aligned(16))) unsigned short sAh[2][MODE == kFP32 ? 8 : AT::LDS_ELEMS];
  __shared__ __attribute__((aligned(16))) float sB[kMT][BN * kLdF];  // [order][col][k]

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, kg = lane >> 4;
  const int wr = wave >> 1, wc = wave & 1;

  for (int t = blockIdx.x; t < w.ntiles; t += gridDim.x) {
    const int4 tile = w.tiles[t];
    if (tile.x < 0) continue;  // idle padding of the XCD-interleaved list
    const GemmGroup& G = w.groups[tile.x];
    const int r0 = tile.y, c0 = tile.z;
    // orders whose rows reach this tile
    int mlive = 0;
    for (int z = 0; z < G.mcount; ++z)
      if (r0 < G.R[z]) mlive = z + 1;
    if (mlive == 0) continue;
    // B tile of one k-group: 32 cols x 32 latitudes x 8 orders (1 KB per column)
    // = 2048 runs of 4 floats, 8 per thread.
    f32x4 bv[8];
    auto load_b = [&](int s_, int k_) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int run = threadIdx.x + 256 * u;  // col = run / 64, within col: 64 runs
        const int col = run >> 6;
        const int r = run & 63;                 // latitude k_ + r/2, orders (r&1)*4 .. +3
        bv[u] = f32x4{0.f, 0.f, 0.f, 0.f};
        if (c0 + col < G.N)
          bv[u] = *reinterpret_cast<const f32x4*>(
              G.B[s_][0] + (long long)(c0 + col) * G.bcs[s_] + (long long)k_ * kMT + r * 4);
      }
    };
    auto store_b = [&]() {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int run = threadIdx.x + 256 * u;
        const int col = run >> 6;
        const int r = run & 63;
        const int i = r >> 1, z0 = (r & 1) * 4;
#pragma unroll
        for (int q = 0; q < 4; ++q) sB[z0 + q][col * kLdF + i] = bv[u][q];
      }
    };
    AT at;
    // step cursor: (sg, k0, z) with z over live orders
    int sg = 0, k0 = 0, z = 0;
    auto advance = [&](int& s_, int& k_, int& z_, bool& newk) {
      newk = false;
      if (++z_ < mlive) return true;
      z_ = 0;
      newk = true;
      k_ += kBK;
      if (k_ < G.Kmax[s_]) return true;
      k_ = 0;
      return ++s_ < G.nseg;
    };

    f32x4 acc[kMT][2], acc2[kMT][2];
#pragma unroll
    for (int a = 0; a < kMT; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
        acc2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
      }

    // prologue
    load_b(sg, k0);
    at.load(G.A[sg][z], G.lda[sg][z], G.piece, r0, G.R[z], k0);
    store_b();
    at.store(sAf[0], sAh[0]);
    __syncthreads();

    int buf = 0;
    bool more = true;
    while (more) {
      int ns = sg, nk = k0, nz = z;
      bool newk;
      const bool has_next = advance(ns, nk, nz, newk);
      if (has_next) {
        at.load(G.A[ns][nz], G.lda[ns][nz], G.piece, r0, G.R[nz], nk);
        if (newk) load_b(ns, nk);
      }
      {
        const float* sa = sAf[buf];
        const unsigned short* sh = sAh[buf];
        const float* sb = sB[z];
        float ax[2][8], bx[8];
        u16x8 ap[2][3], bp[3];
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) read_a<MODE, BM>(sa, sh, wr * 32 + rb * 16 + li, kg, ax[rb], ap[rb]);
        read_b(sb, wc * 16 + li, kg, bx);
        if constexpr (MODE != kFP32) split_pieces<NP>(bx, bp);
#pragma unroll
        for (int zz = 0; zz < kMT; ++zz) {
          if (zz != z) continue;
#pragma unroll
          for (int rb = 0; rb < 2; ++rb) mfma_tile<MODE>(acc[zz][rb], acc2[zz][rb], ax[rb], ap[rb], bx, bp);
        }
      }
      if (has_next) {
        if (newk) {
          __syncthreads();  // every wavefront is done with this k-group's B tile
          store_b();
        }
        at.store(sAf[buf ^ 1], sAh[buf ^ 1]);
      }
      __syncthreads();
      buf ^= 1;
      sg = ns;
      k0 = nk;
      z = nz;
      more = has_next;
    }
    // store every live order's rows
#pragma unroll
    for (int zz = 0; zz < kMT; ++zz) {
      if (zz >= mlive) continue;
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        const int col = c0 + wc * 16 + li;
        const int row = r0 + wr * 32 + rb * 16 + 4 * kg;
        if (col >= G.N || row >= G.R[zz]) continue;
        f32x4 v = acc[zz][rb];
        if constexpr (MODE != kFP32) v = v + acc2[zz][rb];
        *reinterpret_cast<f32x4*>(G.C[zz] + (long long)col * G.ccs + row) = v;
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
#define SHT_LAUNCH(K, M) hipLaunchKernelGGL((sht::K<M>), g, b, 0, s, w)
  if (dir == 0) {
    if (mode == sht::kFP32) SHT_LAUNCH(legendre_syn_kernel, sht::kFP32);
    else if (mode == sht::kBF16X3) SHT_LAUNCH(legendre_syn_kernel, sht::kBF16X3);
    else if (mode == sht::kBF16X6) SHT_LAUNCH(legendre_syn_kernel, sht::kBF16X6);
    else return hipErrorInvalidValue;
  } else {
    if (mode == sht::kFP32) SHT_LAUNCH(legendre_ana_kernel, sht::kFP32);
    else if (mode == sht::kBF16X3) SHT_LAUNCH(legendre_ana_kernel, sht::kBF16X3);
    else if (mode == sht::kBF16X6) SHT_LAUNCH(legendre_ana_kernel, sht::kBF16X6);
    else return hipErrorInvalidValue;
  }
#undef SHT_LAUNCH
  return hipGetLastError();
}

extern "C" int sht_gemm_tile_rows(int dir) { return dir == 0 ? sht::kSynBM : sht::kAnaBM; }
extern "C" int sht_gemm_tile_cols(int dir) { return dir == 0 ? sht::kSynBN : sht::kAnaBN; }
