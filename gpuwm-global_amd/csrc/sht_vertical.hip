// The semi-implicit vertical-mode operators on spectral columns (reference
// semi_implicit.py: VerticalModeSemiImplicit.linear_tendencies :796-813,
// half_apply :845-878, apply :887-919 and solve_shifted :921-959, with
// _finish :815-843).
//
// Every (n, m) column couples the L divergence levels and the L+1
// thermodynamic rows x = [theta, ln ps] through the constant vertical
// matrices B (L x L+1), C (L+1 x L), M (L x L) and the per-degree Helmholtz
// inverse H_n (L x L):
//     forcing = D + k2_n (s_bx B x + s_md M D)
//     d       = H_n forcing
//     x'      = x + s_c C (s_cd D + s_cn d)
// then degree zero passes through (d(0,0) = D(0,0), x'(0,0) = x(0,0)) and
// both are projected (zero above the triangle, m = 0 real).  solve_shifted
// is s_bx = s_c = tau w, s_md = s_cd = 0, s_cn = 1; apply is s_bx = s_c = dt,
// s_md = -alpha (1 - alpha) dt^2, s_cd = 1 - alpha, s_cn = alpha.
//
// Per column the work is ~3 L^2 multiply-adds against 16 L bytes of HBM
// traffic.  Two kernels: vertical_rows_kernel (solve, apply, tendencies)
// keeps each lane's stage column in registers, and vertical_kernel (the half
// map, whose four per-degree matrices need the LDS that the former gives its
// columns) reads both operands of every multiply-add from LDS.
#include "sht_kernels.h"

namespace sht {

// One block per (degree n, COLS orders): the matrices and the block's
// columns sit in LDS; thread (col, grp) owns column col and rows grp,
// grp + G, ... (G = 256 / COLS).  Modes:
//   0: forcing = D + k2 (s_bx B x + s_md M D); d = H_n forcing;
//      x' = x + s_c C (s_cd D + s_cn d); degree zero pinned (apply,
//      solve_shifted)
//   1: d = s_bx k2 B x; x' = s_c C D (linear_tendencies)
//   2: d = P_D D + Q_D (k2 B x); x' = P_x x + Q_x (C D); degree zero pinned
//      (half_apply)
// and every output projected (zero above the triangle, m = 0 real).
template <int COLS>
__global__ void __launch_bounds__(256) vertical_kernel(VerticalArgs a) {
  constexpr int G = 256 / COLS;
  extern __shared__ float smem[];
  const int L = a.nlev, L1 = a.nlev + 1, T1 = a.T + 1;
  const int n = blockIdx.x;
  const int m0 = blockIdx.y * COLS;
  const int tid = threadIdx.x;
  const int col = tid % COLS, grp = tid / COLS;
  const int m = m0 + col;
  const bool live = m < T1;
  const size_t plane = (size_t)T1 * T1;
  const size_t at = (size_t)n * T1 + m;  // (n, m) within a plane

  if (m0 > n) {
    // every order of the block is above the triangle: zeros
    if (live)
      for (int i = grp; i < L1; i += G) {
        if (i < L) a.d_out[(size_t)i * plane + at] = make_float2(0.f, 0.f);
        if (i < L) a.theta_out[(size_t)i * plane + at] = make_float2(0.f, 0.f);
        else a.lnps_out[at] = make_float2(0.f, 0.f);
      }
    return;
  }

  // LDS: B, C, the columns D and x, then per mode
  //   0: M, H_n, forcing/d;  2: P_D, Q_D, P_x, Q_x, l_D, l_x
  float* sB = smem;           // L x L1
  float* sC = sB + L * L1;    // L1 x L
  float* sE = sC + L1 * L;    // mode-dependent matrices
  float* sM = sE;             // L x L (mode 0 with M only)
  float* sH = a.m_matrix ? sM + L * L : sM;  // L x L
  float* sPD = sE;            // L x L
  float* sQD = sPD + L * L;   // L x L
  float* sPX = sQD + L * L;   // L1 x L1
  float* sQX = sPX + L1 * L1; // L1 x L1
  const int mat = a.mode == 0   ? (a.m_matrix ? 2 : 1) * L * L
                  : a.mode == 2 ? 2 * L * L + 2 * L1 * L1
                                : 0;
  float2* sD = reinterpret_cast<float2*>(sE + mat);  // [level][col]
  float2* sX = sD + L * COLS;                         // [row][col]
  float2* sF = sX + L1 * COLS;                        // mode 0: forcing, then d; mode 2: l_D
  float2* sLX = sF + L * COLS;                        // mode 2: l_x
  for (int e = tid; e < L * L1; e += 256) {
    sB[e] = a.b_matrix[e];
    sC[e] = a.c_matrix[e];
  }
  if (a.mode == 0) {
    const float* Hn = a.inverse + (size_t)n * L * L;
    for (int e = tid; e < L * L; e += 256) {
      if (a.m_matrix) sM[e] = a.m_matrix[e];
      sH[e] = Hn[e];
    }
  } else if (a.mode == 2) {
    const size_t dn = (size_t)n * L * L, xn = (size_t)n * L1 * L1;
    for (int e = tid; e < L * L; e += 256) {
      sPD[e] = a.p_d[dn + e];
      sQD[e] = a.q_d[dn + e];
    }
    for (int e = tid; e < L1 * L1; e += 256) {
      sPX[e] = a.p_x[xn + e];
      sQX[e] = a.q_x[xn + e];
    }
  }
  const bool in_tri = live && m <= n;
  for (int i = grp; i < L1; i += G) {
    float2 v = make_float2(0.f, 0.f), w = make_float2(0.f, 0.f);
    if (in_tri) {
      if (i < L) v = a.div[(size_t)i * plane + at];
      w = i < L ? a.theta[(size_t)i * plane + at] : a.lnps[at];
    }
    if (i < L) sD[i * COLS + col] = v;
    sX[i * COLS + col] = w;
  }
  __syncthreads();

  // row i of (A v) for an LDS matrix A (row stride ld) and an LDS column set v
  auto dot = [&](const float* A, int ld, int i, const float2* v, int len) {
    float2 acc = make_float2(0.f, 0.f);
    for (int j = 0; j < len; ++j) {
      const float c = A[i * ld + j];
      const float2 x = v[j * COLS + col];
      acc.x = fmaf(c, x.x, acc.x);
      acc.y = fmaf(c, x.y, acc.y);
    }
    return acc;
  };
  const float k2 = a.k2[n];
  const bool origin = n == 0 && m == 0;
  const bool pin = origin && a.mode != 1;
  auto store = [&](int i, float2 d, float2 xo) {
    // d for i < L, x' for every i; projected
    if (!live) return;
    if (!in_tri) d = xo = make_float2(0.f, 0.f);
    if (m == 0) d.y = xo.y = 0.f;
    if (i < L) {
      a.d_out[(size_t)i * plane + at] = d;
      a.theta_out[(size_t)i * plane + at] = xo;
    } else {
      a.lnps_out[at] = xo;
    }
  };

  if (a.mode == 1) {
    for (int i = grp; i < L1; i += G) {
      float2 d = make_float2(0.f, 0.f);
      if (i < L) {
        const float2 bx = dot(sB, L1, i, sX, L1);
        d = make_float2(a.s_bx * k2 * bx.x, a.s_bx * k2 * bx.y);
      }
      const float2 cd = dot(sC, L, i, sD, L);
      store(i, d, make_float2(a.s_c * cd.x, a.s_c * cd.y));
    }
    return;
  }
  if (a.mode == 2) {
    // l_D = k2 B x, l_x = C D
    for (int i = grp; i < L1; i += G) {
      if (i < L) {
        const float2 bx = dot(sB, L1, i, sX, L1);
        sF[i * COLS + col] = make_float2(k2 * bx.x, k2 * bx.y);
      }
      sLX[i * COLS + col] = dot(sC, L, i, sD, L);
    }
    __syncthreads();
    for (int i = grp; i < L1; i += G) {
      float2 d = make_float2(0.f, 0.f);
      if (i < L) {
        const float2 pd = dot(sPD, L, i, sD, L), qd = dot(sQD, L, i, sF, L);
        d = pin ? sD[i * COLS + col] : make_float2(pd.x + qd.x, pd.y + qd.y);
      }
      const float2 px = dot(sPX, L1, i, sX, L1), qx = dot(sQX, L1, i, sLX, L1);
      const float2 xo = pin ? sX[i * COLS + col] : make_float2(px.x + qx.x, px.y + qx.y);
      store(i, d, xo);
    }
    return;
  }
  // mode 0: forcing = D + k2 (s_bx B x + s_md M D)
  for (int i = grp; i < L; i += G) {
    const float2 bx = dot(sB, L1, i, sX, L1);
    const float2 md = a.m_matrix ? dot(sM, L, i, sD, L) : make_float2(0.f, 0.f);
    const float2 d = sD[i * COLS + col];
    sF[i * COLS + col] = make_float2(d.x + k2 * (a.s_bx * bx.x + a.s_md * md.x),
                                     d.y + k2 * (a.s_bx * bx.y + a.s_md * md.y));
  }
  __syncthreads();
  // d = H_n forcing (in registers until every thread has read sF)
  float2 dn[(kMaxLevels + G - 1) / G];
  int q = 0;
  for (int i = grp; i < L; i += G, ++q) dn[q] = dot(sH, L, i, sF, L);
  __syncthreads();
  q = 0;
  for (int i = grp; i < L; i += G, ++q) sF[i * COLS + col] = dn[q];
  __syncthreads();
  // x' = x + s_c C (s_cd D + s_cn d)
  for (int i = grp; i < L1; i += G) {
    float2 acc = make_float2(0.f, 0.f);
    for (int j = 0; j < L; ++j) {
      const float c = sC[i * L + j];
      const float2 d = sD[j * COLS + col];
      const float2 e = sF[j * COLS + col];
      const float2 u = make_float2(a.s_cd * d.x + a.s_cn * e.x, a.s_cd * d.y + a.s_cn * e.y);
      acc.x = fmaf(c, u.x, acc.x);
      acc.y = fmaf(c, u.y, acc.y);
    }
    const float2 x = sX[i * COLS + col];
    const float2 xo = pin ? x : make_float2(x.x + a.s_c * acc.x, x.y + a.s_c * acc.y);
    const float2 d = i < L ? (pin ? sD[i * COLS + col] : sF[i * COLS + col])
                           : make_float2(0.f, 0.f);
    store(i, d, xo);
  }
}

// Solve, apply and tendencies (modes 0 and 1): one block of four wavefronts
// per (degree n, 64 orders m), lane = order.  The block's matrices sit in LDS
// (row stride LP, zero padded); wavefront w computes the rows i = w, w + 4,
// ... of each stage from the lane's column of the stage's right-hand side,
// held in registers (v[j], zero from its length on), and reads each group of
// four matrix values with one 16-byte load at a wave-uniform address.  LDS
// then carries 4 bytes per lane and multiply-add pair, not 12, and passes
// each stage's result column (64 orders wide) on to the next stage.
template <int LP>
__device__ __forceinline__ float2 lds_row_dot(const float* row, const float2 (&v)[LP]) {
  typedef __attribute__((ext_vector_type(4))) float f4;
  float ax = 0.f, ay = 0.f;
#pragma unroll
  for (int q = 0; q < LP / 4; ++q) {
    const f4 c = *reinterpret_cast<const f4*>(row + 4 * q);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      ax = fmaf(c[k], v[4 * q + k].x, ax);
      ay = fmaf(c[k], v[4 * q + k].y, ay);
    }
  }
  return make_float2(ax, ay);
}

template <int LP>
__global__ void __launch_bounds__(256) vertical_rows_kernel(VerticalArgs a) {
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  extern __shared__ float smem[];
  const int L = a.nlev, L1 = a.nlev + 1, T1 = a.T + 1;
  const int n = blockIdx.x;
  const int m0 = blockIdx.y * 64;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int m = m0 + lane;
  const bool live = m < T1;
  const bool in_tri = live && m <= n;
  const unsigned plane8 = (unsigned)T1 * (unsigned)T1 * 8u;  // bytes per level plane
  // Columns through raw buffers: a level's value at (n, m) is plane row i
  // (scalar offset i x plane8) plus the lane's byte offset; lanes off the
  // triangle (loads) or past T (stores) take an offset beyond every buffer,
  // so their loads return zeros and their stores are dropped, with no
  // per-lane address registers beyond one offset.
  constexpr unsigned kOff = 0x80000000u;
  const unsigned at8 = ((unsigned)n * (unsigned)T1 + (unsigned)m) * 8u;
  const unsigned ld_off = in_tri ? at8 : kOff, st_off = live ? at8 : kOff;
  auto rsrc = [](const void* p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000);
  };
  const auto rD = rsrc(a.div, (unsigned)L * plane8), rT = rsrc(a.theta, (unsigned)L * plane8);
  const auto rP = rsrc(a.lnps, plane8);
  const auto rDo = rsrc(a.d_out, (unsigned)L * plane8), rTo = rsrc(a.theta_out, (unsigned)L * plane8);
  const auto rPo = rsrc(a.lnps_out, plane8);
  auto load = [&](__amdgpu_buffer_rsrc_t r, int i) {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(
                                          r, (int)ld_off, (int)((unsigned)i * plane8), 0));
  };
  auto store = [&](__amdgpu_buffer_rsrc_t r, int i, float2 v) {  // projected (m = 0 real)
    if (m == 0) v.y = 0.f;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)st_off,
                                          (int)((unsigned)i * plane8), 0);
  };
  auto D = [&](int i) { return load(rD, i); };
  auto X = [&](int i) { return i < L ? load(rT, i) : load(rP, 0); };
  auto store_x = [&](int i, float2 v) {
    if (i < L) store(rTo, i, v);
    else store(rPo, 0, v);
  };
  const float2 zero = make_float2(0.f, 0.f);
  if (m0 > n) {  // every order of the block is above the triangle: zeros
    for (int i = w; i < L1; i += 4) {
      if (i < L) store(rDo, i, zero);
      store_x(i, zero);
    }
    return;
  }
  // LDS: B (L rows), C (L1 rows), [mode 0: M (L rows, with M), H_n (L rows)],
  // row stride LP; then one result column [L][64]
  float* sB = smem;
  float* sC = sB + (size_t)L * LP;
  float* sM = sC + (size_t)L1 * LP;
  float* sH = a.m_matrix ? sM + (size_t)L * LP : sM;
  float2* scol = reinterpret_cast<float2*>(a.mode == 0 ? sH + (size_t)L * LP : sM);
  auto stage = [&](float* dst, const float* src, int rows, int len) {
    for (int r = w; r < rows; r += 4)
      for (int j = lane; j < LP; j += 64) dst[(size_t)r * LP + j] = j < len ? src[(size_t)r * len + j] : 0.f;
  };
  stage(sB, a.b_matrix, L, L1);
  stage(sC, a.c_matrix, L1, L);
  if (a.mode == 0) {
    if (a.m_matrix) stage(sM, a.m_matrix, L, L);
    stage(sH, a.inverse + (size_t)n * L * L, L, L);
  }
  auto S = [&](int i) -> float2& { return scol[(size_t)i * 64 + lane]; };
  float2 v[LP];
  auto load_v = [&](auto&& src, int len) {
#pragma unroll
    for (int j = 0; j < LP; ++j) v[j] = j < len ? src(j) : zero;
  };
  const float k2 = a.k2[n];
  const bool pin = n == 0 && m == 0 && a.mode == 0;
  __syncthreads();  // the matrices are in LDS

  if (a.mode == 1) {
    // d = s_bx k2 B x; x' = s_c C D
    load_v(X, L1);
    for (int i = w; i < L; i += 4) {
      const float2 bx = lds_row_dot<LP>(sB + (size_t)i * LP, v);
      store(rDo, i, make_float2(a.s_bx * k2 * bx.x, a.s_bx * k2 * bx.y));
    }
    load_v(D, L);
    for (int i = w; i < L1; i += 4) {
      const float2 cd = lds_row_dot<LP>(sC + (size_t)i * LP, v);
      store_x(i, make_float2(a.s_c * cd.x, a.s_c * cd.y));
    }
    return;
  }
  // A row's own D or x value is loaded one row ahead (the next row of the
  // wavefront, i + 4), so that its latency hides behind the current row.
  // forcing = D + k2 (s_bx B x + s_md M D)
  load_v(X, L1);
  float2 dn = w < L ? D(w) : zero;
  for (int i = w; i < L; i += 4) {
    const float2 d = dn;
    if (i + 4 < L) dn = D(i + 4);
    const float2 bx = lds_row_dot<LP>(sB + (size_t)i * LP, v);
    if (a.m_matrix) S(i) = bx;
    else S(i) = make_float2(d.x + k2 * a.s_bx * bx.x, d.y + k2 * a.s_bx * bx.y);
  }
  if (a.m_matrix) {  // the rows this wavefront wrote above
    load_v(D, L);
    dn = w < L ? D(w) : zero;
    for (int i = w; i < L; i += 4) {
      const float2 d = dn;
      if (i + 4 < L) dn = D(i + 4);
      const float2 md = lds_row_dot<LP>(sM + (size_t)i * LP, v);
      const float2 bx = S(i);
      S(i) = make_float2(d.x + k2 * (a.s_bx * bx.x + a.s_md * md.x),
                         d.y + k2 * (a.s_bx * bx.y + a.s_md * md.y));
    }
  }
  __syncthreads();  // the whole forcing column is in LDS
  // d = H_n forcing; the column then becomes u = s_cd D + s_cn d
  load_v(S, L);
  __syncthreads();  // every wavefront holds the forcing before it is overwritten
  dn = w < L ? D(w) : zero;
  for (int i = w; i < L; i += 4) {
    const float2 di = dn;
    if (i + 4 < L) dn = D(i + 4);
    const float2 dd = lds_row_dot<LP>(sH + (size_t)i * LP, v);
    store(rDo, i, pin ? di : dd);
    S(i) = make_float2(a.s_cd * di.x + a.s_cn * dd.x, a.s_cd * di.y + a.s_cn * dd.y);
  }
  __syncthreads();
  // x' = x + s_c C u
  load_v(S, L);
  float2 xn = w < L1 ? X(w) : zero;
  for (int i = w; i < L1; i += 4) {
    const float2 x = xn;
    if (i + 4 < L1) xn = X(i + 4);
    const float2 cu = lds_row_dot<LP>(sC + (size_t)i * LP, v);
    store_x(i, pin ? x : make_float2(x.x + a.s_c * cu.x, x.y + a.s_c * cu.y));
  }
}



}  // namespace sht

// Orders per block: 32, or 16 for the half map, whose four per-degree
// matrices would not leave LDS for 32 columns at 64 levels.
static int vertical_cols(int mode) { return mode == 2 ? 16 : 32; }

static size_t vertical_lds_bytes(int nlev, int mode, bool with_m) {
  const size_t L = nlev, L1 = nlev + 1, cols = vertical_cols(mode);
  const size_t mat = mode == 0 ? (with_m ? 2 : 1) * L * L : mode == 2 ? 2 * L * L + 2 * L1 * L1 : 0;
  const size_t colsets = mode == 0 ? L + L1 + L : mode == 2 ? L + L1 + L + L1 : L + L1;
  return (2 * L * L1 + mat) * sizeof(float) + colsets * cols * sizeof(float2);
}

// LDS of vertical_rows_kernel: the staged matrices at row stride lp and one
// result column of 64 orders.
static size_t vertical_rows_lds_bytes(int nlev, int mode, bool with_m, int lp) {
  const size_t L = nlev, L1 = nlev + 1;
  const size_t rows = L + L1 + (mode == 0 ? (with_m ? 2 : 1) * L : 0);
  return rows * lp * sizeof(float) + L * 64 * sizeof(float2);
}

extern "C" hipError_t sht_launch_vertical_solve(const sht::VerticalArgs* a, hipStream_t s) {
  if (a->nlev < 1 || a->nlev > sht::kMaxLevels || a->mode < 0 || a->mode > 2)
    return hipErrorInvalidValue;
  // vertical_rows_kernel for solve, apply and tendencies while its buffer
  // offsets (32-bit, beyond 2^31 for the lanes it masks) cover the planes
  const size_t planes = (size_t)a->nlev * (size_t)(a->T + 1) * (size_t)(a->T + 1) * 8;
  if (a->mode != 2 && planes < ((size_t)1 << 31)) {
    // the stage columns in registers: nlev + 1 values, rounded up to groups of 8
    const int lp = (a->nlev + 1 + 7) / 8 * 8;
    const size_t lds = vertical_rows_lds_bytes(a->nlev, a->mode, a->m_matrix != nullptr, lp);
    const void* k = nullptr;
    switch (lp) {
#define SHT_VK(LP) \
  case LP: k = reinterpret_cast<const void*>(sht::vertical_rows_kernel<LP>); break;
      SHT_VK(8) SHT_VK(16) SHT_VK(24) SHT_VK(32) SHT_VK(40) SHT_VK(48) SHT_VK(56) SHT_VK(64)
      SHT_VK(72)
#undef SHT_VK
      default: return hipErrorInvalidValue;
    }
    hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const dim3 g((unsigned)(a->T + 1), (unsigned)((a->T + 1 + 63) / 64)), b(256);
    sht::VerticalArgs copy = *a;
    void* args[] = {&copy};
    return hipLaunchKernel(k, g, b, args, lds, s);
  }
  const int cols = vertical_cols(a->mode);
  const size_t lds = vertical_lds_bytes(a->nlev, a->mode, a->m_matrix != nullptr);
  const void* k = cols == 16 ? reinterpret_cast<const void*>(sht::vertical_kernel<16>)
                             : reinterpret_cast<const void*>(sht::vertical_kernel<32>);
  hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  const dim3 g((unsigned)(a->T + 1), (unsigned)((a->T + 1 + cols - 1) / cols));
  if (cols == 16) hipLaunchKernelGGL(sht::vertical_kernel<16>, g, dim3(256), lds, s, *a);
  else hipLaunchKernelGGL(sht::vertical_kernel<32>, g, dim3(256), lds, s, *a);
  return hipGetLastError();
}
