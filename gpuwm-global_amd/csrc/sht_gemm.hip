// Grouped Legendre GEMM over the triangular m-schedule (one persistent launch
// per direction).
//
// Every Legendre contraction of a batch is a list of work items: a panel (a
// group of up to kMT = 4 consecutive orders m of one job and one output
// parity, times one 64-row tile) and a first column (BN = 32, 64 or 96
// columns per item; 96 has no split accumulators).  Per order m the group holds
//     C_m[r][c] = sum_seg sum_k A_seg,m[r][k] * B_seg,m[k][c]
// with A a Legendre table block and B a data block; two segments express the
// vector operators (P*m/cos and dP/dphi against differently arranged data) as
// one accumulation.  The host lists each panel's k-steps (one order, 32 k) in
// execution order, with byte addresses and valid extents, so the kernel does
// no schedule arithmetic: panels come in through the scalar cache, step
// descriptors and tiles through bounds-checked buffer loads (reads past an
// extent return 0, which is also how row and column edges are handled).
//
// A step may read its B columns through a map (reverse each group of four,
// negate some of them): the vector operators' second arrangement of the data
// is such a map of the first, so only one copy is stored.
//
// Block tile 64 x BN, four wavefronts of 16 rows x BN (16x16 MFMA blocks).
// Each wavefront loads its own table fragments (16 rows x 32 k per piece)
// from the tiled table straight into registers: no table tile passes through
// LDS.  Only the data tile B, which all four wavefronts read, is staged in
// LDS (SYN: two buffers, one per step; ANA: one k-chunk, resident for its
// trip of four steps).
//
// Pipeline: four register stages (step j's tiles sit in stage j % 4, its
// order within the trip).  While the wavefronts run step j's MFMAs, the
// loads of steps j+1, j+2 and j+3 are in flight (the loop is bound by memory
// latency otherwise).  SYN writes step j+1's B to the other LDS buffer and
// one barrier ends the step; ANA needs barriers only where its resident B
// tile changes.  The panel's step descriptors sit in LDS for the whole tile,
// so no memory load other than the tiles is waited on inside the loop.
// Reads past the panel's list return its last null step (extent 0), so every
// load is issued unconditionally.
//
// LDS B tiles are unpadded rows of 32 k, one row per column (128 bytes fp32,
// 64 bytes per bf16 piece), whose 16-byte chunks are XOR-swizzled by row, so
// that the MFMA fragment reads (16 columns x one chunk) and the tile writes
// are free of bank conflicts (measured).
//
// SYN (synthesis): rows = folded latitudes, k = degrees, C = waist
//   W[c][m/4][p][i][m%4].  Each lane keeps four order accumulators (4
//   latitudes x 4 orders = 64 contiguous bytes of the waist); the tile is
//   transposed through LDS so that stores write whole column runs.
// ANA (analysis): rows = degrees, k = folded latitudes.  The waist puts the
//   four orders of one latitude side by side, so one B tile (32 latitudes x
//   BN columns x 4 orders) is loaded once per k-chunk and serves the steps
//   of all four orders; steps flagged with a B extent start a new k-chunk.
//
// Precision modes:
//   FP32   v_mfma_f32_16x16x4_f32, fp32 tables and data.
//   BF16X3 v_mfma_f32_16x16x32_bf16 on hi+lo pieces: a0b0 + a0b1 + a1b0.
//   BF16X6 three pieces each: a0b0 + (a0b1 + a1b0) + (a0b2 + a1b1 + a2b0),
//          every term down to 2^-16 relative.  With ACC2 the five small
//          products accumulate apart from a0b0 and are added at the end.
// Tables are split once on the device from FP64 (sht_tables.hip); the data
// stay fp32 in memory and are split as their tile is written to LDS.  Each
// piece is the round-to-nearest-even bf16 of what the previous pieces leave.
// Lane l of a
// 16x16 MFMA: (l & 15) is the row/column and (l >> 4) selects the k chunk;
// A and B use the same k mapping.
#include <type_traits>

#include "sht_kernels.h"

namespace sht {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef __attribute__((ext_vector_type(4))) unsigned int u32x4;
typedef __attribute__((ext_vector_type(2))) unsigned int u32x2;

// Address spaces of the device pass (the host pass only parses the kernels).
#if defined(__HIP_DEVICE_COMPILE__)
#define SHT_CONST __attribute__((address_space(4)))   // scalar-cache loads
#define SHT_GLOBAL __attribute__((address_space(1)))  // global (not flat) stores
#else
#define SHT_CONST
#define SHT_GLOBAL
#endif

constexpr int kBK = 32;  // k per step
constexpr unsigned kRsrcWord3 = 0x00020000u;  // raw buffer, 32-bit data format

// A step descriptor as loaded by every lane (two 16-byte vector loads); the
// fields are uniform and move to scalar registers with readfirstlane.
struct GemmDesc {
  u32x4 lo, hi;  // {a lo, a hi, b lo, b hi}, {a_bytes, b_klen, z, cmap}
  __device__ __forceinline__ GemmStep scalar() const {
    auto rfl = [](unsigned v) { return (unsigned)__builtin_amdgcn_readfirstlane((int)v); };
    GemmStep s;
    s.a = ((unsigned long long)rfl(lo[1]) << 32) | rfl(lo[0]);
    s.b = ((unsigned long long)rfl(lo[3]) << 32) | rfl(lo[2]);
    s.a_bytes = rfl(hi[0]);
    s.b_klen = (int)rfl(hi[1]);
    s.z = (int)rfl(hi[2]);
    s.cmap = (int)rfl(hi[3]);
    return s;
  }
};

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
__device__ __forceinline__ u32x4 load16(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
}

// 16-byte chunk swizzle of an LDS row: fp32 rows have 8 chunks, bf16 rows 4.
// Chosen by measurement on gfx950 (SQ_LDS_BANK_CONFLICT over the exact
// fragment-read and tile-write patterns of this kernel): both the 16-row x
// one-chunk fragment reads and the whole-row writes are conflict-free.
template <bool F32>
__device__ __forceinline__ int swz(int row, int chunk) {
  if constexpr (F32) return chunk ^ ((row >> 1) & 7);
  else return chunk ^ ((row >> 1) & 3);
}

template <int MODE, int DIR, int BN>
struct GemmCfg {
  static constexpr bool F32 = MODE == kFP32;
  static constexpr int NP = F32 ? 1 : (MODE == kBF16X6 ? 3 : 2);  // stored pieces
  static constexpr int ESZ = F32 ? 4 : 2;   // element bytes of tables and data
  static constexpr int ROWB = kBK * ESZ;    // bytes of one row of a 32-k tile
  static constexpr int BZ = DIR == 0 ? 1 : kMT;
  static constexpr int NB = BN / 16;        // 16-column MFMA blocks per wavefront
  static constexpr int B_TILE = BZ * BN * ROWB;        // LDS bytes per piece
  static constexpr int B_BUFS = DIR == 0 ? 2 : 1;      // ANA keeps one k-chunk resident
  // register stage, in 16-byte registers per lane: the wavefront's own A
  // fragments (16 rows x 32 k: one register per bf16 piece, two for fp32),
  // loaded straight from the tiled table; the thread's share of the B tile,
  // fp32 (split into pieces when written to LDS).
  static constexpr int A_REGS = F32 ? 2 : NP;
  static constexpr int B_REGS = BZ * BN * kBK * 4 / (256 * 16);
  static_assert(B_REGS >= 1, "tile shape");
};

// Four fp32 values as `np` bf16 piece vectors (each the round-to-nearest-even
// bf16 of what the previous pieces leave; the residuals are exact in fp32).
// Pairs are converted together (v_cvt_pk_bf16_f32); a packed pair's values
// return to fp32 as its low half shifted up and its high half masked.
__device__ __forceinline__ void split4(const float (&x)[4], int np, u32x2 (&out)[3]) {
  typedef __attribute__((ext_vector_type(2))) float f2;
  typedef __attribute__((ext_vector_type(2))) __bf16 b2;
  f2 r[2] = {f2{x[0], x[1]}, f2{x[2], x[3]}};
#pragma unroll
  for (int pc = 0; pc < 3; ++pc) {
    if (pc >= np) break;
    unsigned w[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      w[h] = __builtin_bit_cast(unsigned, __builtin_convertvector(r[h], b2));
      if (pc + 1 < np) {
        r[h][0] -= __uint_as_float(w[h] << 16);
        r[h][1] -= __uint_as_float(w[h] & 0xFFFF0000u);
      }
    }
    out[pc] = u32x2{w[0], w[1]};
  }
}

template <int MODE>
__device__ __forceinline__ void mfma_step(f32x4& acc, f32x4& acc2, const u32x4 (&a)[3],
                                          const u32x4 (&b)[3]) {
#define SHT_MFMA_BF(ACC, X, Y)                                                                 \
  ACC = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, X),                \
                                                __builtin_bit_cast(bf16x8, Y), ACC, 0, 0, 0)
  if constexpr (MODE == kFP32) {
    // a[0], b[0] hold k chunk kg, a[1], b[1] chunk kg + 4 (4 floats each)
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int s = 0; s < 4; ++s)
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(a[h][s]),
                                                   __uint_as_float(b[h][s]), acc, 0, 0, 0);
  } else {
    if constexpr (MODE == kBF16X6) {
      SHT_MFMA_BF(acc2, a[2], b[0]);
      SHT_MFMA_BF(acc2, a[1], b[1]);
      SHT_MFMA_BF(acc2, a[0], b[2]);
    }
    SHT_MFMA_BF(acc2, a[1], b[0]);
    SHT_MFMA_BF(acc2, a[0], b[1]);
    SHT_MFMA_BF(acc, a[0], b[0]);
  }
#undef SHT_MFMA_BF
}

// 32-column tiles fit three workgroups per CU (<= 168 VGPRs), 64- and
// 96-column ones two (the 96-column analysis tile's LDS, B tile plus step
// list, is exactly half of the CU's 160 KB in bf16x6).
template <int MODE, int DIR, int BN, bool ACC2>
__global__ void __launch_bounds__(256, BN == 32 ? 3 : 2) legendre_gemm_kernel(GemmWork w) {
  using C = GemmCfg<MODE, DIR, BN>;
  constexpr bool F32 = C::F32;
  constexpr int NP = C::NP, ROWB = C::ROWB, NB = C::NB;
  __shared__ __attribute__((aligned(16))) unsigned char sB[C::B_BUFS][NP * C::B_TILE];
  __shared__ u32x4 sDesc[2 * kGemmMaxSteps];  // the current panel's step list

  const SHT_CONST GemmPanel* panels = (const SHT_CONST GemmPanel*)w.panels;
  const SHT_CONST int2* tiles = (const SHT_CONST int2*)w.tiles;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, kg = lane >> 4;
  const int arow = 16 * wave + li;  // the row of this lane's A fragment

  for (int t = blockIdx.x; t < w.ntiles; t += gridDim.x) {
    const int2 tile = tiles[t];
    if (tile.x < 0) continue;
    const GemmPanel P = panels[tile.x];
    const int c0 = tile.y;
    const int ncol = min(BN, P.N - c0);
    const unsigned long long coloff = (unsigned long long)c0 * P.bcs * 4;  // fp32 data
    const int step0 = P.step0, nsteps = P.nsteps;

    // ---- register stages and their loads: step j's tiles sit in stage
    // j % 4 (its order within the trip), so loads run three steps ahead.
    // ANA loads B once per trip (one stage). ----
    constexpr int kStages = kMT;
    constexpr int kBStages = DIR == 0 ? kStages : 1;
    u32x4 ra[kStages][C::A_REGS], rb[kBStages][C::B_REGS];
    // The panel's step list (with its two null steps) is copied to LDS once
    // per tile; a step's descriptor is then an LDS broadcast read, off the
    // memory queue that the tile loads wait in.
    {
      const __amdgpu_buffer_rsrc_t srs = make_rsrc(
          reinterpret_cast<unsigned long long>(w.steps) + (unsigned long long)step0 * 32,
          (unsigned)(nsteps + 2) * 32u);
      for (int e = tid; e < 2 * (nsteps + 2); e += 256) sDesc[e] = load16(srs, (unsigned)e * 16u);
      __syncthreads();
    }
    auto fetch_desc = [&](int j) {
      const int e = 2 * min(j, nsteps + 1);
      return GemmDesc{sDesc[e], sDesc[e + 1]};
    };
    // WITH_B: the step may carry a B tile (SYN: every step; ANA: the first
    // step of each k-chunk, position 0 of a trip of kMT steps)
    auto issue = [&](u32x4(&A)[C::A_REGS], u32x4(&B)[C::B_REGS], const GemmStep& s,
                     auto WITH_B) {
      // A: this lane's fragment of each piece, 16 bytes (bf16: k chunk kg of
      // row arow; fp32: chunks kg and kg + 4); rows past a_bytes read 0
#pragma unroll
      for (int pc = 0; pc < NP; ++pc) {
        const __amdgpu_buffer_rsrc_t r =
            make_rsrc(s.a + (unsigned long long)(pc * P.piece_bytes), s.a_bytes);
        if constexpr (F32) {
          A[0] = load16(r, (unsigned)(arow * ROWB + kg * 16));
          A[1] = load16(r, (unsigned)(arow * ROWB + (kg + 4) * 16));
        } else {
          A[pc] = load16(r, (unsigned)(arow * ROWB + kg * 16));
        }
      }
      const int rev = (s.cmap & 1) ? 3 : 0;
      if constexpr (DIR == 0) {
        // packed fp32 columns, k contiguous: a thread takes 4 k of one column
        const unsigned bytes =
            s.b_klen > 0 ? (unsigned)(((ncol - 1) * P.bcs + s.b_klen) * 4) : 0u;
        const __amdgpu_buffer_rsrc_t r = make_rsrc(s.b + coloff, bytes);
#pragma unroll
        for (int u = 0; u < C::B_REGS; ++u) {
          const int q = tid + 256 * u;
          const int col = q >> 3;
          B[u] = load16(r, (unsigned)((col ^ rev) * P.bcs * 4 + (q & 7) * 16));
        }
      } else if constexpr (decltype(WITH_B)::value) {
        // k-chunk of the fp32 waist: per column 32 latitudes x 4 orders; a
        // thread takes 4 latitudes (a quad, 64 bytes) of one column
        const unsigned bytes =
            s.b_klen > 0 ? (unsigned)(((ncol - 1) * P.bcs + s.b_klen * kMT) * 4) : 0u;
        const __amdgpu_buffer_rsrc_t r = make_rsrc(s.b + coloff, bytes);
#pragma unroll
        for (int u = 0; u < BN / 32; ++u) {
          const int q = tid + 256 * u;
          const int col = q >> 3, quad = q & 7;
          const unsigned off = (unsigned)((col ^ rev) * P.bcs * 4 + quad * 64);
#pragma unroll
          for (int h = 0; h < 4; ++h) B[u * 4 + h] = load16(r, off + 16u * h);
        }
      }
    };
    // B stage -> LDS (split into pieces in the bf16 modes); `cmap` bits 1..4
    // negate columns 0..3 of each group of four
    auto put_b = [&](const u32x4(&B)[C::B_REGS], int buf, int cmap) {
      auto flip_of = [&](int col) { return ((cmap >> (1 + (col & 3))) & 1) ? 0x80000000u : 0u; };
      if constexpr (DIR == 0) {
#pragma unroll
        for (int u = 0; u < C::B_REGS; ++u) {
          const int q = tid + 256 * u;
          const int col = q >> 3, ch = q & 7;  // 4 k starting at 4 ch
          const u32x4 v = B[u] ^ flip_of(col);
          unsigned char* row = &sB[buf][col * ROWB];
          if constexpr (F32) {
            *reinterpret_cast<u32x4*>(row + swz<true>(col, ch) * 16) = v;
          } else {
            const float x[4] = {__uint_as_float(v[0]), __uint_as_float(v[1]),
                                __uint_as_float(v[2]), __uint_as_float(v[3])};
            u32x2 pcs[3];
            split4(x, NP, pcs);
#pragma unroll
            for (int pc = 0; pc < NP; ++pc)
              *reinterpret_cast<u32x2*>(row + pc * C::B_TILE + swz<false>(col, ch >> 1) * 16 +
                                        (ch & 1) * 8) = pcs[pc];
          }
        }
      } else {
        (void)buf;
#pragma unroll
        for (int u = 0; u < BN / 32; ++u) {
          const int q = tid + 256 * u;
          const int col = q >> 3, quad = q & 7;
          const unsigned flip = flip_of(col);
          const u32x4* v = &B[u * 4];  // v[h] = latitude 4 quad + h, four orders
#pragma unroll
          for (int z = 0; z < kMT; ++z) {
            unsigned char* row = &sB[0][(z * BN + col) * ROWB];
            const u32x4 wv = u32x4{v[0][z], v[1][z], v[2][z], v[3][z]} ^ flip;
            if constexpr (F32) {
              *reinterpret_cast<u32x4*>(row + swz<true>(col, quad) * 16) = wv;
            } else {
              const float x[4] = {__uint_as_float(wv[0]), __uint_as_float(wv[1]),
                                  __uint_as_float(wv[2]), __uint_as_float(wv[3])};
              u32x2 pcs[3];
              split4(x, NP, pcs);
#pragma unroll
              for (int pc = 0; pc < NP; ++pc)
                *reinterpret_cast<u32x2*>(row + pc * C::B_TILE + swz<false>(col, quad >> 1) * 16 +
                                          (quad & 1) * 8) = pcs[pc];
            }
          }
        }
      }
    };

    f32x4 acc[kMT][NB], acc2[kMT][NB];
#pragma unroll
    for (int a = 0; a < kMT; ++a)
#pragma unroll
      for (int n = 0; n < NB; ++n) {
        acc[a][n] = f32x4{0.f, 0.f, 0.f, 0.f};
        acc2[a][n] = f32x4{0.f, 0.f, 0.f, 0.f};
      }

    // one step into the accumulators of order Z: A fragments from register
    // stage Z, B fragments from LDS buffer `buf`
    auto compute = [&](int buf, auto ZC) {
      constexpr int Z = decltype(ZC)::value;
      const int bbuf = DIR == 0 ? buf : 0;
      constexpr int bz = DIR == 0 ? 0 : Z;
      u32x4 af[3];
#pragma unroll
      for (int r = 0; r < C::A_REGS; ++r) af[r] = ra[Z][r];
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const int col = nb * 16 + li;
        const unsigned char* base = &sB[bbuf][(bz * BN + col) * ROWB];
        u32x4 bf[3];
        if constexpr (F32) {
#pragma unroll
          for (int h = 0; h < 2; ++h)
            bf[h] = *reinterpret_cast<const u32x4*>(base + swz<true>(col, kg + 4 * h) * 16);
        } else {
#pragma unroll
          for (int pc = 0; pc < NP; ++pc)
            bf[pc] =
                *reinterpret_cast<const u32x4*>(base + pc * C::B_TILE + swz<false>(col, kg) * 16);
        }
        if constexpr (ACC2) mfma_step<MODE>(acc[Z][nb], acc2[Z][nb], af, bf);
        else mfma_step<MODE>(acc[Z][nb], acc[Z][nb], af, bf);
      }
    };

    // ---- prologue: steps 0, 1 and 2 in flight, step 0's B in LDS buffer 0 ----
    // q1, q2, q3: steps j+1, j+2, j+3 (scalar) at the start of step j; dn:
    // the raw descriptor of step j+4, fetched one step ahead.
    GemmStep q1, q2, q3;
    GemmDesc dn;
    {
      const GemmStep s0 = fetch_desc(0).scalar();
      q1 = fetch_desc(1).scalar();
      q2 = fetch_desc(2).scalar();
      issue(ra[0], rb[0], s0, std::true_type{});
      issue(ra[1], rb[DIR == 0 ? 1 : 0], q1, std::integral_constant<bool, DIR == 0>{});
      issue(ra[2], rb[DIR == 0 ? 2 : 0], q2, std::integral_constant<bool, DIR == 0>{});
      q3 = fetch_desc(3).scalar();
      dn = fetch_desc(4);
      put_b(rb[0], 0, s0.cmap);
    }
    __syncthreads();

    // ---- main loop: trips of kMT steps, one per order Z of the panel (the
    // host pads every k-chunk to kMT steps), so each step's accumulators,
    // register stage and LDS buffer are fixed at compile time ----
    auto body = [&](auto ZC, int j) {
      constexpr int Z = decltype(ZC)::value;
      constexpr int s = Z & 1;               // LDS buffer of step j's B (SYN)
      constexpr int R3 = (Z + 3) % kStages;  // register stage (= trip position) of step j+3
      constexpr int R1 = (Z + 1) % kStages;  // register stage of step j+1
      issue(ra[R3], rb[DIR == 0 ? R3 : 0], q3, std::integral_constant<bool, DIR == 0 || R3 == 0>{});
      const GemmStep q4 = dn.scalar();
      dn = fetch_desc(j + 5);
      compute(s, ZC);
      if constexpr (DIR == 0) {
        // step j+1's B goes to the other buffer; the barrier orders it before
        // step j+1 reads it and after every read of this buffer two steps ago
        if (j + 1 < nsteps) put_b(rb[R1], s ^ 1, q1.cmap);
        __syncthreads();
      } else if constexpr (R1 == 0) {
        // ANA: the resident B tile changes with the next trip
        __syncthreads();  // every wavefront is done with it
        if (j + 1 < nsteps) put_b(rb[0], 0, q1.cmap);
        __syncthreads();
      }
      q1 = q2;
      q2 = q3;
      q3 = q4;
    };
    static_assert(kMT == 4, "trip of four steps");
    for (int j = 0; j < nsteps; j += kMT) {
      body(std::integral_constant<int, 0>{}, j);
      body(std::integral_constant<int, 1>{}, j + 1);
      body(std::integral_constant<int, 2>{}, j + 2);
      body(std::integral_constant<int, 3>{}, j + 3);
    }

    // ---- epilogue: this lane holds rows 16 wave + 4 kg + q (q < 4) of
    // column 16 nb + li ----
    auto wave_lds_sync = [] {  // this wavefront's LDS accesses, in order
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    if constexpr (DIR == 0) {
      // A lane's four rows x four orders are 64 contiguous bytes of the
      // waist, and the four lanes of a column 256.  Each column block passes
      // through the wavefront's own LDS area, transposed, so that a store
      // instruction writes four whole 256-byte column runs rather than 16-byte
      // pieces of sixteen columns (measured: ~10 us of a T383 synthesis).
      // each wavefront's 16 rows x 16 columns of 4-order values (rows padded
      // to 17 against bank conflicts); declared here, so the analysis
      // instantiations keep their LDS budget
      __shared__ f32x4 sOut[4 * 16 * 17];
      f32x4* mine = sOut + wave * 16 * 17;
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          f32x4 o;
#pragma unroll
          for (int zz = 0; zz < kMT; ++zz) {
            f32x4 v = acc[zz][nb];
            if constexpr (!F32 && ACC2) v = v + acc2[zz][nb];
            o[zz] = v[q];
          }
          mine[li * 17 + 4 * kg + q] = o;
        }
        wave_lds_sync();
        const int r = lane & 15, row = 16 * wave + r;
#pragma unroll
        for (int tt = 0; tt < 4; ++tt) {  // columns 4 tt + lane / 16 of the block, row r
          const int cl = 4 * tt + (lane >> 4), col = nb * 16 + cl;
          const f32x4 o = mine[cl * 17 + r];
          if (col < ncol && row < P.rows[0])
            *reinterpret_cast<SHT_GLOBAL f32x4*>(reinterpret_cast<SHT_GLOBAL float*>(P.c[0]) +
                                                 (long long)(c0 + col) * P.ccs +
                                                 (long long)row * kMT) = o;
        }
        wave_lds_sync();  // the area is read before the next block lands
      }
    } else {
      // rows row..row+3 of one order are 16 contiguous bytes of the packed
      // output, a column's 64 rows (four wavefronts) 256: per order, the
      // block's tile passes through LDS (the B tile's space, free after the
      // loop) so that a store instruction writes four whole column runs
      // (measured: FP32 analysis with 32-column tiles 187 -> 171 us at T383)
      constexpr int CS = 68;  // floats per staged column: 64 rows + pad
      static_assert(BN * CS * 4 <= C::B_BUFS * NP * C::B_TILE, "staging fits the B tile");
      float* stage = reinterpret_cast<float*>(&sB[0][0]);
#pragma unroll
      for (int zz = 0; zz < kMT; ++zz) {
        __syncthreads();  // the B tile (or the previous order's stage) is consumed
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
          f32x4 v = acc[zz][nb];
          if constexpr (!F32 && ACC2) v = v + acc2[zz][nb];
          *reinterpret_cast<f32x4*>(stage + (nb * 16 + li) * CS + 16 * wave + 4 * kg) = v;
        }
        __syncthreads();
        const int r4 = 4 * (lane & 15);
#pragma unroll
        for (int tt = 0; tt < BN / 16; ++tt) {  // columns 4 (wave + 4 tt) + lane / 16
          const int col = 4 * (wave + 4 * tt) + (lane >> 4);
          const f32x4 o = *reinterpret_cast<const f32x4*>(stage + col * CS + r4);
          if (col < ncol && r4 < P.rows[zz])
            *reinterpret_cast<SHT_GLOBAL f32x4*>(reinterpret_cast<SHT_GLOBAL float*>(P.c[zz]) +
                                                 (long long)(c0 + col) * P.ccs + r4) = o;
        }
      }
    }
    __syncthreads();  // sB and sDesc are reused by the next tile
  }
}

}  // namespace sht

namespace {

template <int MODE, int DIR>
hipError_t launch_mode(const sht::GemmWork& w, int bn, int acc2, dim3 g, hipStream_t s) {
  const dim3 b(256);
  constexpr bool kSplit = MODE != sht::kFP32;
  if (bn == 96) {
    // no split accumulators: 96 columns leave no registers for them
    hipLaunchKernelGGL((sht::legendre_gemm_kernel<MODE, DIR, 96, false>), g, b, 0, s, w);
  } else if (bn == 64) {
    if (kSplit && acc2)
      hipLaunchKernelGGL((sht::legendre_gemm_kernel<MODE, DIR, 64, kSplit>), g, b, 0, s, w);
    else
      hipLaunchKernelGGL((sht::legendre_gemm_kernel<MODE, DIR, 64, false>), g, b, 0, s, w);
  } else {
    if (kSplit && acc2)
      hipLaunchKernelGGL((sht::legendre_gemm_kernel<MODE, DIR, 32, kSplit>), g, b, 0, s, w);
    else
      hipLaunchKernelGGL((sht::legendre_gemm_kernel<MODE, DIR, 32, false>), g, b, 0, s, w);
  }
  return hipGetLastError();
}

template <int DIR>
hipError_t launch_dir(int mode, const sht::GemmWork& w, int bn, int acc2, dim3 g,
                      hipStream_t s) {
  switch (mode) {
    case sht::kFP32: return launch_mode<sht::kFP32, DIR>(w, bn, acc2, g, s);
    case sht::kBF16X3: return launch_mode<sht::kBF16X3, DIR>(w, bn, acc2, g, s);
    case sht::kBF16X6: return launch_mode<sht::kBF16X6, DIR>(w, bn, acc2, g, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

extern "C" hipError_t sht_launch_legendre_gemm(const sht::GemmLaunch* L, hipStream_t s) {
  if (L->ntiles <= 0) return hipSuccess;
  if (L->bn != 32 && L->bn != 64 && L->bn != 96) return hipErrorInvalidValue;
  if (L->bn == 96 && L->acc2 && L->mode != sht::kFP32) return hipErrorInvalidValue;
  const sht::GemmWork w{L->panels, L->steps, L->tiles, L->ntiles};
  const dim3 g(L->grid);
  return L->dir == 0 ? launch_dir<0>(L->mode, w, L->bn, L->acc2, g, s)
                     : launch_dir<1>(L->mode, w, L->bn, L->acc2, g, s);
}
