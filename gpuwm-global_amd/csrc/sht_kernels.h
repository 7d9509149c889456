// Kernel argument structs and launchers shared by the HIP kernels and the
// host-side C ABI (sht_capi.cpp).
#pragma once

#include "sht_common.h"

namespace sht {

constexpr int kMT = 4;           // m-tile of the waist layout W[c][m/4][p][i][m%4]
constexpr int kFftThreads = 256;
constexpr int kMaxStages = 16;

// ---- tables (sht_tables.hip) ----
// The block of order m and parity p starts at koff[2m+p] * JHP elements, holds
// JHP folded latitudes i x KP = kp(m,p) degrees (n - m)/2, both multiples of
// 32, and is stored in 32-wide k-chunks so that one GEMM k-step of any 64 rows
// is one contiguous run:
//   synthesis (rows i, k = degree):   chunk k/32, then i, then k%32
//   analysis  (rows degree, k = i):   chunk i/32, then degree, then i%32
__host__ __device__ inline long long syn_tiled(int JHP, int i, int nidx) {
  return (long long)(nidx >> 5) * JHP * 32 + (long long)i * 32 + (nidx & 31);
}
__host__ __device__ inline long long ana_tiled(int KP, int i, int nidx) {
  return (long long)(i >> 5) * KP * 32 + (long long)nidx * 32 + (i & 31);
}

struct TableOut {
  // One entry per kind (0 = P, 1 = dP/dphi, 2 = P*m/cos), NULL to skip.
  float* f32_syn[3];
  float* f32_ana[3];
  unsigned short* bf_syn[3];  // three pieces back to back, piece stride = `elems`
  unsigned short* bf_ana[3];
  double* f64_syn[3];         // optional FP64 copy (tests)
  long long elems;            // elements of one table in one layout
  // FP64 tables of the orders m < low_m0 for the vector operators' FP64
  // correction, over r = n - m (T+1 slots) and folded rows i (JHP):
  //   low_syn [m][r][i]   low_ana [m][i][r]
  // except the basis (kind 0) over degrees m..T+1: low_syn[0] [m][r < T+2][i],
  // low_ana[0] [m][i][r < T+2]
  double* f64_low_syn[3];
  double* f64_low_ana[3];
  int low_m0;
};

// ---- FFT (sht_fft.hip) ----
struct FftPlan {
  int N;                      // nlon
  int wave_L;                 // N / 64 when the wave-resident path runs, else 0
  int nstages;
  int radix[kMaxStages];
  const float2* tw;           // tw[k] = exp(-2 pi i k / N), k < N
  const double2* tw64;        // the same in FP64 (the FP64 FFTs)
};

struct FoldJob {
  const float* grid0;         // field rows at grid0 + f*nlat*nlon (u for vector jobs)
  const float* grid1;         // v rows for vector jobs, else NULL
  float* waist0;              // scalar: waist; vector: [iV, iU] per level (fp32)
  const float* rowscale;      // per folded row: 1/nlon x quadrature weight (x 1/a)
  const float* rowscale_hi;   // wave-resident fold: orders m >= m_hi take this scale
  int m_hi;                   // (the recurrence form's 1/cos(phi)); NULL: rowscale
  int nfields;                // scalar fields, or u/v level pairs
  int col0;                   // first data column of this job in the waist
  int ldw;                    // floats per data column
  int vector;                 // 0 scalar, 1 vector pair
  const float* goff;          // scalar jobs: per field, subtracted from its rows (or NULL)
  // FP64 fold of one field of a vector pair (rfft_fold_wave64_kernel):
  // writes i X to columns col0 + fstride f + {0, 1} of the fp32 waist and,
  // for orders m < m0, the folded (E, O) of X in FP64 to low + f low_stride,
  // laid out [m][p][i]
  double2* low;
  const double* rowscale64;   // per folded row, FP64
  const double* rowscale64_hi;  // orders m >= m0 of the fp32 waist (NULL: rowscale64)
  int fstride;
  int m0;
  int low_stride;             // double2 per field
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
  const float* rowscale_hi;   // wave-resident unfold: orders m >= m_hi take this scale
  int m_hi;                   // (the recurrence form's 1/cos(phi)); NULL: rowscale
  int nfields;
  int col0;                   // columns of field f: col0 + fstride*f + {0 re, 1 im}
  int fstride;
  int ldw;
  int imul;                   // multiply the spectrum by 1: i, 2: i m before the FFT
  int post_exp;               // write exp() of the grid values
  // FP64 unfold (irfft_unfold_wave64_kernel): the spectrum is assembled and
  // inverted in FP64, orders m < m0 read from low + f low_stride ([m][p][i]
  // folded (E, O) in FP64) instead of the fp32 waist
  int f64;
  const double2* low;
  const double* rowscale64;
  const double* rowscale64_hi;  // orders m >= m0 (NULL: rowscale64)
  int m0;
  int low_stride;
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

// ---- the reference's Fourier waist (*lead, nlat, T+1) complex (sht_fft.hip) ----
// Rows <-> waist rows: the real FFT along longitude of `nrows` rows per lead
// index, divided by nlon and cut to m <= T (rows -> waist), or its inverse
// (waist -> rows: irfft of nlon x the retained orders).
struct WaistRowsArgs {
  FftPlan fft;
  int T, nlat;
  const float* rows_in;       // (nlead, nrows, nlon) rows -> waist
  float* rows_out;            // (nlead, nrows, nlon) waist -> rows
  const float2* waist_in;     // (nlead, nlat, T+1)
  float2* waist_out;
  int nlead, nrows, row0;     // waist rows row0 .. row0 + nrows of each lead index
};
// Reference waist -> the analysis GEMM's folded waist W[c][m/4][p][i][m%4]
// (parity sums of the north/south row pair times the quadrature weight).
struct WaistFoldArgs {
  const float2* w0;           // (nfields, nlat, T+1): scalar fields, or u
  const float2* w1;           // v (vector jobs)
  float* waist;
  float* waist1;              // vector == 2: the second waist (-V)
  const float* rowscale;      // per folded row: 2 pi w (x 1/a for vector jobs)
  int T, nlat, JH, JHP, ldw, nfields, col0;
  int vector;                 // 0 scalar; 1 [iV, iU] per level; 2 iU to waist, -V to waist1
};
// The synthesis GEMM's folded waist -> reference waist (north = E + O,
// south = E - O).
struct WaistUnfoldArgs {
  const float* waist;
  float2* out;                // (nfields, nlat, T+1)
  int T, nlat, JH, JHP, ldw, nfields, col0, fstride;
};

// ---- point sampler (sht_sample.hip) ----
struct SamplePackArgs {
  const float2* in0;  // (count, T+1, T+1): the fields, or vorticity for winds
  const float2* in1;  // divergence for winds
  void* out;          // (K, ldc) complex128 (f64) or complex64, K = (T+1)(T+2)/2 in triangle order
  double radius;
  int T, count, ldc, wind, f64;
};
constexpr int kSampleMaxParts = 8;
struct SampleCuts {
  int parts;                      // R: order-range parts per (point block, field tile)
  int m[kSampleMaxParts + 1];     // part r holds orders m[r] .. m[r+1]-1
};
struct SampleArgs {
  const void* coef;       // (K, ldc) from the pack
  const double* rec;      // [3][K]: recurrence a, b and derivative c per packed (n, m)
  const double* diag;     // [2][T+1]: -sqrt((2m+1)/2m), sqrt(2m+3)
  const double* sin_lat;  // (npts)
  const double* cos_lat;  // (npts), for the east metric
  const double* lon;      // (npts) radians
  double* out0;           // (nout, npts)
  double* out1;
  double* partial;        // R > 1: (2 R, ldc, npts) raw sums of the parts
  SampleCuts cuts;
  double radius;
  int T, npts, ldc, nout, kind, f32;
};

// ---- semi-implicit vertical-mode solves (sht_vertical.hip) ----
constexpr int kMaxLevels = 64;
struct VerticalArgs {
  const float2* div;      // (L, T+1, T+1)
  const float2* theta;    // (L, T+1, T+1)
  const float2* lnps;     // (T+1, T+1)
  const float* b_matrix;  // (L, L+1)
  const float* c_matrix;  // (L+1, L)
  const float* m_matrix;  // (L, L), NULL when s_md = 0
  const float* inverse;   // (T+1, L, L), mode 0
  const float* p_d;       // mode 2: (T+1, L, L) P_D, Q_D and (T+1, L+1, L+1) P_x, Q_x
  const float* q_d;
  const float* p_x;
  const float* q_x;
  const float* k2;        // (T+1)
  float2* d_out;
  float2* theta_out;
  float2* lnps_out;
  float s_bx, s_md, s_c, s_cd, s_cn;
  int T, nlev, mode;      // 0 solve/apply, 1 linear tendencies, 2 half map
};

// ---- grouped GEMM (sht_gemm.hip) ----
// A panel is one row tile of one work group: up to kMT consecutive orders m
// of one job and one output parity.  Its k-steps (one order, 32 k each) are
// listed on the host, in the order the kernel runs them; the kernel reads
// panels and steps through the scalar cache.  Work items are (panel, first
// column).  All addresses are byte addresses.
struct GemmStep {
  unsigned long long a;       // A tile: rows r0.. of the panel, this k-chunk (tiled table)
  unsigned long long b;       // B at column 0: SYN packed block + k0; ANA waist block + k0
  unsigned int a_bytes;       // valid bytes from a (rows past the order's end read 0)
  int b_klen;                 // valid k from b (SYN: K - k0; ANA: JHP - k0); 0 = no B load
  int z;                      // order within the m-tile
  int cmap;                   // B column map: 0 = as stored; else bit 0 reverses each
                              // group of 4 columns and bit 1+c negates column c of a group
};

constexpr int kGemmBM = 64;      // GEMM block tile rows
constexpr int kGemmMaxSteps = 256;  // steps per panel incl. the two null steps (LDS-resident):
                                     // two segments of JHP/32 chunks x 4 orders, JHP <= 992;
                                     // 8 KB, which with the 72 KB B tile of the 96-column
                                     // bf16x6 analysis leaves room for two blocks per CU

struct GemmPanel {
  unsigned long long c[kMT];  // SYN: c[0] = waist block + r0 rows; ANA: packed block of order z + r0
  long long piece_bytes;      // bf16 piece stride of the tables (data B are fp32)
  int step0, nsteps;          // steps[step0 .. step0+nsteps), followed by two null steps
  int N;                      // data columns
  int bcs;                    // B column stride, floats
  int ccs;                    // C column stride, floats
  int rows[kMT];              // rows of C written per order (<= tile rows)
  int pad;
};
static_assert(sizeof(GemmStep) == 32, "GemmStep is one 8-dword scalar load");

// One grouped-GEMM launch: precision mode, direction (0 SYN, 1 ANA), column
// tile (32 or 64), separate accumulators for the small split products.
struct GemmLaunch {
  int mode, dir, bn, acc2;
  const GemmPanel* panels;
  const GemmStep* steps;
  const int2* tiles;
  int ntiles, grid;
};

// ---- the vector operators' FP64 low orders (sht_lowm.hip) ----
// Orders m < m0 of uv_from_vordiv and vordiv_from_uv run in FP64 end to end.
// The low waist of a vector job is [level][f][m][p][i] double2 (complex),
// f = 0 for u and 1 for v, p the parity, i the folded row (JHP slots).
struct LowSynArgs {
  const float2* vort;         // (nlev, T+1, T+1) [n][m]
  const float2* div;
  const double* p;            // low_syn[0]: the basis [m][r < T+2][i]
  double2* low;               // (E, O) of a cos(phi) (u, v): the unfold applies 1/(a cos)
  double radius;
  int T, JH, JHP, m0, nlev;
};
struct LowAnaArgs {
  const double2* low;         // (E, O) of u and v from the FP64 fold, weights
                              // 2 pi w / (a nlon cos(phi)) applied
  const double* p;            // low_ana[0]: the basis [m][i][r < T+2]
  double2* ab;                // A, B of the orders m < m0: [level][A|B][n < T+2][m < m0]
  float2* vort;               // (nlev, T+1, T+1): entries n >= m, m < m0
  float2* div;
  int T, JH, JHP, m0, nlev;
};

// ---- spectral data movement (sht_pack.hip) ----
enum PackMode : int { kPlain = 0, kIM = 1, kICoef = 2, kVector = 3, kInvLap = 4, kVectorRec = 5 };

struct PackJob {
  const float2* in0;          // [f][n][m]; vorticity for VECTOR
  const float2* in1;          // divergence for VECTOR
  float* out0;                // packed ([-psi, chi] per level for VECTOR; fp32)
  int nfields;                // fields, or levels for VECTOR
  int col0;
  int ldp;                    // floats per packed column
  int mode;
  float radius;
  const float2* off00;        // kPlain: per field, added to its (0, 0) coefficient (or NULL)
  int m0;                     // kVectorRec: orders m < m0 are not written
  int mdp;                    // kVectorRec: orders m < mdp as [-psi, chi] (two-segment form)
};

struct PackArgs {
  const int* koff;            // koff[2m+p]
  int T;
  int mtiles;                 // ceil((T+1)/32): 32-order tiles
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
  int m0;                     // orders m < m0 inside the triangle are left alone
                              // (written by the FP64 low-order kernel)
  int rec;                    // vector jobs in the recurrence form: columns [iB, iA]
                              // over degrees m..T+1 (see unpack_vector_rec_kernel)
  int mdp;                    // rec: orders m < mdp hold [zeta, D] (two-segment form)
  const float2* off00;        // scalar jobs: per field, added to its (0, 0) output (or NULL)
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
hipError_t sht_fft_allow_lds(size_t bytes, size_t wave_bytes);
int sht_fft_wave_L(int N);                       // 0: no wave-resident path for N
size_t sht_fft_wave_lds_bytes(int N, int vector);
hipError_t sht_launch_legendre_gemm(const sht::GemmLaunch* launch, hipStream_t s);
hipError_t sht_launch_waist_from_rows(const sht::WaistRowsArgs* a, hipStream_t s);
hipError_t sht_launch_rows_from_waist(const sht::WaistRowsArgs* a, hipStream_t s);
hipError_t sht_launch_waist_fold(const sht::WaistFoldArgs* a, hipStream_t s);
hipError_t sht_launch_waist_unfold(const sht::WaistUnfoldArgs* a, hipStream_t s);
hipError_t sht_launch_sample_pack(const sht::SamplePackArgs* a, hipStream_t s);
hipError_t sht_launch_sample(const sht::SampleArgs* a, hipStream_t s);
hipError_t sht_launch_vertical_solve(const sht::VerticalArgs* a, hipStream_t s);
hipError_t sht_launch_pack(const sht::PackArgs* a, hipStream_t s);
hipError_t sht_launch_low_synthesis(const sht::LowSynArgs* a, hipStream_t s);
hipError_t sht_launch_low_analysis(const sht::LowAnaArgs* a, hipStream_t s);
hipError_t sht_low_allow_lds(size_t bytes);
size_t sht_fold64_lds_bytes(int N);
size_t sht_unfold64_lds_bytes(int N);
hipError_t sht_launch_unpack(const sht::UnpackArgs* a, hipStream_t s);
hipError_t sht_launch_spectral_elementwise(const float2* in, float2* out, long long total, int T,
                                           int op, float radius, hipStream_t s);
}
