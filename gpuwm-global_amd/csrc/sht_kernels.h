// Kernel argument structs and launchers shared by the HIP kernels and the
// host-side C ABI (sht_capi.cpp).
#pragma once

#include "sht_common.h"

namespace sht {

constexpr int kMT = 8;           // m-tile of the waist layout W[c][m/8][p][i][m%8]
constexpr int kFftThreads = 256;
constexpr int kMaxStages = 16;

// ---- tables (sht_tables.hip) ----
struct TableOut {
  // One entry per kind (0 = P, 1 = dP/dphi, 2 = P*m/cos), NULL to skip.
  float* f32_syn[3];
  float* f32_ana[3];
  unsigned short* bf_syn[3];  // three pieces back to back, piece stride = `elems`
  unsigned short* bf_ana[3];
  double* f64_syn[3];         // optional FP64 copy (tests)
  long long elems;            // elements of one table in one layout
};

// ---- FFT (sht_fft.hip) ----
struct FftPlan {
  int N;                      // nlon
  int nstages;
  int radix[kMaxStages];
  const float2* tw;           // tw[k] = exp(-2 pi i k / N), k < N
};

struct FoldJob {
  const float* grid0;         // field rows at grid0 + f*nlat*nlon (u for vector jobs)
  const float* grid1;         // v rows for vector jobs, else NULL
  float* waist0;              // scalar: waist; vector: arrangement a = [iV, iU]
  float* waist1;              // vector: arrangement b = [U, -V]
  const float* rowscale;      // per folded row: 1/nlon x quadrature weight (x 1/a)
  int nfields;                // scalar fields, or u/v level pairs
  int col0;                   // first data column of this job in the waist
  int ldw;                    // floats per data column
  int vector;                 // 0 scalar, 1 vector pair
};

struct FoldArgs {
  FftPlan fft;
  int T, nlat, JH, JHP;
  int rp;                     // folded rows (wavefronts) per block (1, 2 or 4)
  int ngroups;                // folded-row groups covering JHP rows
  int gpb;                    // groups per block
  int nchunks;                // ceil(ngroups / gpb)
  FoldJob job;
};

struct UnfoldJob {
  const float* waist;         // folded E/O waist
  float* grid;                // field f rows at grid + f*nlat*nlon
  const float* rowscale;      // per folded row multiplier
  int nfields;
  int col0;                   // columns of field f: col0 + fstride*f + {0 re, 1 im}
  int fstride;
  int ldw;
  int imul;                   // multiply the spectrum by i before the FFT
};

struct UnfoldArgs {
  FftPlan fft;
  int T, nlat, JH, JHP;
  int rp;
  int ngroups;                // ceil(JH / rp)
  int gpb;                    // groups per block
  int nchunks;                // ceil(ngroups / gpb)
  UnfoldJob job;
};

// ---- grouped GEMM (sht_gemm.hip) ----
// One work group: up to kMT consecutive orders m of one job and one output
// parity.  Per order: C_m = sum_seg A_seg,m x B_seg,m (see sht_gemm.hip).
struct GemmGroup {
  const void* A[2][kMT];      // table block of (seg, m): A[r][k] at A + r*lda + k
  const float* B[2][kMT];     // SYN: packed block of (seg, m); ANA: waist block (m-tile, parity)
  float* C[kMT];              // SYN: C[0] = waist block (m-tile, parity); ANA: packed block of m
  long long piece;            // bf16 piece stride (elements) of the tables
  int K[2][kMT];              // k extent per (seg, m), multiple of 32; 0 = no work
  int lda[2][kMT];
  int R[kMT];                 // rows per m (multiple of 32)
  int Kmax[2];
  int bks[2], bcs[2];         // B strides (k, column); ANA reads the waist's m-tiled k
  int nseg, N, Rmax, mcount;  // segments, columns, max rows, orders in the tile
  int ccs;                    // C column stride
};

// ---- spectral data movement (sht_pack.hip) ----
enum PackMode : int { kPlain = 0, kIM = 1, kICoef = 2, kVector = 3 };

struct PackJob {
  const float2* in0;          // [f][n][m]; vorticity for VECTOR
  const float2* in1;          // divergence for VECTOR
  float* out0;                // packed (arrangement a for VECTOR)
  float* out1;                // arrangement b for VECTOR
  int nfields;                // fields, or levels for VECTOR
  int col0;
  int ldp;                    // floats per packed column
  int mode;
  float radius;
};

struct PackArgs {
  const int* koff;            // koff[2m+p]
  int T;
  int mtiles, ntiles;         // ceil((T+1)/32), ceil(KP(0,0)/32)
  PackJob job;
};

struct UnpackJob {
  const float* in;
  float2* out0;               // [f][n][m] (vorticity for vector jobs)
  float2* out1;               // divergence for vector jobs
  int nfields;
  int col0;
  int ldp;
  int vector;                 // columns per level [zeta.re, zeta.im, D.re, D.im]
};

struct UnpackArgs {
  const int* koff;
  int T;
  int tiles;                  // ceil((T+1)/32)
  UnpackJob job;
};

}  // namespace sht

extern "C" {
hipError_t sht_launch_gauss_newton(int nlat, double* mu, double* w, hipStream_t s);
hipError_t sht_launch_legendre_tables(int T, int nlat, int JH, int JHP, const double* mu,
                                      const int* koff, const sht::TableOut* out, hipStream_t s);
hipError_t sht_launch_rfft_fold(const sht::FoldArgs* a, hipStream_t s);
hipError_t sht_launch_irfft_unfold(const sht::UnfoldArgs* a, hipStream_t s);
size_t sht_fold_lds_bytes(int N, int T, int vector, int rp);
size_t sht_unfold_lds_bytes(int N, int T, int rp);
hipError_t sht_fft_allow_lds(size_t bytes);
hipError_t sht_launch_legendre_gemm(int mode, int dir, const sht::GemmGroup* groups,
                                    const int4* tiles, int ntiles, int grid, hipStream_t s);
int sht_gemm_tile_rows(int dir);
int sht_gemm_tile_cols(int dir);
hipError_t sht_launch_pack(const sht::PackArgs* a, hipStream_t s);
hipError_t sht_launch_unpack(const sht::UnpackArgs* a, hipStream_t s);
hipError_t sht_launch_spectral_elementwise(const float2* in, float2* out, long long total, int T,
                                           int op, float radius, hipStream_t s);
}
