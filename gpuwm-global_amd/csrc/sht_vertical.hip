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
// One block per (degree n, 32 orders): the matrices and the block's columns
// sit in LDS, each thread owns one column and a strided eighth of the
// levels, and every stage is a short fp32 dot product per output (the work
// is ~3 L^2 multiply-adds per column against 16 L bytes of HBM traffic, so
// the kernel is bound by HBM, not by arithmetic).
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

extern "C" hipError_t sht_launch_vertical_solve(const sht::VerticalArgs* a, hipStream_t s) {
  if (a->nlev < 1 || a->nlev > sht::kMaxLevels || a->mode < 0 || a->mode > 2)
    return hipErrorInvalidValue;
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
