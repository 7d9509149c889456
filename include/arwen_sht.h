/*
 * arwen_sht.h -- C ABI of the MI355X spherical-harmonic transform engine.
 *
 * Drop-in boundary for the reference's transform layer
 * (src/arwen_global/spectral/transform.py, vector.py).  Every data pointer
 * is DEVICE memory owned by the caller, in the reference's array layouts:
 *
 *   grid field      (*lead, nlat, nlon)   float32, C-contiguous, latitudes
 *                                         ascending (south to north)
 *   spectral field  (*lead, T+1, T+1)     complex64 (interleaved re, im),
 *                                         indexed [n, m], zero above the
 *                                         triangle on output, m = 0 real
 *                                         after analysis/projection
 *
 * "count" is the flattened lead size (fields x levels x members).  Every
 * operation is enqueued on `stream` (a hipStream_t; NULL = the null stream)
 * and returns without synchronizing.  Functions return ARWEN_SHT_OK or an
 * error code; arwen_sht_last_error() gives this thread's last message.
 * No exception crosses the ABI.  A plan belongs to the device that was
 * current when it was created; call operations with that device current.
 */
#ifndef ARWEN_SHT_H
#define ARWEN_SHT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define ARWEN_SHT_OK 0
#define ARWEN_SHT_EINVAL 1 /* bad argument: Python raises ValueError   */
#define ARWEN_SHT_EHIP 2   /* HIP runtime error: Python raises RuntimeError */
#define ARWEN_SHT_ENOMEM 3 /* device allocation failed                 */

/* Legendre GEMM precision. */
#define ARWEN_SHT_FP32 1   /* v_mfma_f32_16x16x4_f32 on fp32 tables            */
#define ARWEN_SHT_BF16X3 2 /* bf16 hi+lo pairs, 3 MFMA products                 */
#define ARWEN_SHT_BF16X6 3 /* bf16 hi+lo+residual pieces, 6 MFMA products       */

/* Operation codes of a batch job (arwen_sht_job.op). */
#define ARWEN_SHT_OP_SYNTHESIS 1           /* in0 coeff -> out0 grid                 */
#define ARWEN_SHT_OP_ZONAL_DERIVATIVE 2    /* in0 coeff -> out0 grid of inverse(c*i*m) */
#define ARWEN_SHT_OP_MERIDIONAL_DERIVATIVE 3 /* in0 coeff -> out0 grid of d/dphi     */
#define ARWEN_SHT_OP_GRADIENT 4            /* in0 coeff -> out0 east, out1 north     */
#define ARWEN_SHT_OP_UV_FROM_VORDIV 5      /* in0 vort, in1 div -> out0 u, out1 v    */
#define ARWEN_SHT_OP_ANALYSIS 6            /* in0 grid -> out0 coeff                 */
#define ARWEN_SHT_OP_VORDIV_FROM_UV 7      /* in0 u, in1 v -> out0 vort, out1 div    */
/* The reference's Fourier waist: (count, nlat, T+1) complex64, the zonal
 * spectrum of each latitude row divided by nlon and cut to m <= T
 * (FourierWaist, reference transform.py:417-629). */
#define ARWEN_SHT_OP_SYNTHESIS_TO_WAIST 8  /* in0 coeff -> out0 waist (basis)        */
#define ARWEN_SHT_OP_ZONAL_TO_WAIST 9      /* in0 coeff -> out0 waist of c*i*m       */
#define ARWEN_SHT_OP_MERIDIONAL_TO_WAIST 10 /* in0 coeff -> out0 waist (dP/dphi)     */
#define ARWEN_SHT_OP_ANALYSIS_FROM_WAIST 11 /* in0 waist -> out0 coeff               */
#define ARWEN_SHT_OP_VORDIV_FROM_WAIST 12  /* in0 u waist, in1 v waist -> vort, div  */
/* inverse, plus the east and north gradients of the same field: out0 grid,
 * out1 east = d/dlambda / (a cos), out2 north = d/dphi / a (reference
 * transform.py:1742-1768).  Two plane contractions per field: the east
 * derivative is the grid's own waist times i m, as inverse(c i m) / (a cos)
 * is in the reference (inverse_zonal_derivative, :1752-1760). */
#define ARWEN_SHT_OP_SYNTHESIS_GRADIENT 13
/* in0 u waist, in1 v waist -> out0 divergence alone: the zonal component
 * against P m / cos and the meridional one against dP/dphi, two plane
 * contractions per level where the vorticity-and-divergence pair takes four
 * (reference vector.py:136-166, divergence_from_fourier). */
#define ARWEN_SHT_OP_DIVERGENCE_FROM_WAIST 14

/* arwen_sht_job.flags */
#define ARWEN_SHT_JOB_EXP 1  /* SYNTHESIS: out0 = exp(inverse(in0)), the surface
                                pressure from ln ps (reference dynamics.py:964-965)
                                or a moisture field kept >= 0 by a log transform */
#define ARWEN_SHT_JOB_INVLAP 2 /* SYNTHESIS: out0 = inverse(inverse_laplacian(in0)),
                                  the streamfunction or velocity potential of a
                                  vorticity or divergence (reference
                                  da/ensemble.py:569-570), fused into the pack */

typedef struct arwen_sht_plan arwen_sht_plan;
typedef struct arwen_sht_batch arwen_sht_batch;

typedef struct arwen_sht_plan_desc {
  int truncation;         /* T >= 1                                          */
  int nlat;               /* Gaussian latitudes, >= T+1                      */
  int nlon;               /* even, >= 2T+2                                   */
  double radius_m;        /* sphere radius (reference constants.py:4)        */
  int precision;          /* ARWEN_SHT_FP32 / _BF16X3 / _BF16X6              */
  const double* sin_lat;  /* optional HOST nodes, ascending; NULL: the device
                             Newton iteration computes them                  */
  const double* weights;  /* HOST quadrature weights (with sin_lat)          */
  int fp64_orders;        /* vector operators (uv_from_vordiv, vordiv_from_uv):
                             zonal orders m < fp64_orders run in FP64 end to
                             end (FP64 longitude FFTs for u and v, FP64
                             tables and sums for those orders); the rest use
                             the plan's precision, as contractions of the
                             basis alone through the derivative recurrence
                             (two per level instead of four; with 0, from
                             order 4 on where nlon = 64 L).  -1: the
                             default, 0 for
                             ARWEN_SHT_FP32 and 4 ceil((T+1)/192) for the
                             split-BF16 modes; 0: off; k > 0: rounded up to
                             a multiple of 4.  Needs nlon = 64 L with
                             L in {2,3,4,6,8,12,16,24}; other grids run with
                             0.  No reference counterpart: the reference's
                             float32 path runs these orders in fp32.        */
} arwen_sht_plan_desc;

/* Build a plan: Gauss nodes and weights (FP64 Newton on the device) and the
 * parity-folded Legendre tables P, dP/dphi and P*m/cos(phi) (FP64 recurrence
 * on the device, stored in the plan's precision).  Replaces
 * SphericalHarmonicTransform.create/__post_init__/_build_tables
 * (reference transform.py:728-786, 837-933, 953-1023) and
 * GaussianGrid.create (grid.py:96-126). */
int arwen_sht_plan_create(const arwen_sht_plan_desc* desc, arwen_sht_plan** out);
int arwen_sht_plan_destroy(arwen_sht_plan* plan);

/* Host copies of the plan's nodes (ascending sin latitude) and weights,
 * nlat each.  Replaces GaussianGrid.sin_lat / quadrature_weights
 * (reference grid.py:112-125). */
int arwen_sht_plan_grid(const arwen_sht_plan* plan, double* sin_lat, double* weights);

/* Gauss-Legendre nodes (ascending sin latitude) and weights of `nlat` points
 * by FP64 Newton iteration on the current device, copied to HOST arrays of
 * nlat doubles.  Replaces numpy.polynomial.legendre.leggauss as called by
 * GaussianGrid.create (reference grid.py:112). */
int arwen_sht_gauss_nodes(int nlat, double* sin_lat, double* weights);

/* Resident device bytes of the plan's Legendre tables.  Replaces
 * legendre_table_nbytes (reference transform.py:935-951). */
size_t arwen_sht_plan_table_bytes(const arwen_sht_plan* plan);

/* The plan's FP64 vector orders (arwen_sht_plan_desc.fp64_orders as
 * resolved: 0 or a multiple of 4). */
int arwen_sht_plan_fp64_orders(const arwen_sht_plan* plan);

/* FP64 Legendre table (kind 0 = P, 1 = dP/dphi, 2 = P*m/cos) for the folded
 * north rows, written to HOST memory `out` as out[m][n][i] (zero for n < m),
 * (T+1) x (T+1) x ceil(nlat/2) doubles.  Exposes the device recurrence that
 * replaces legendre_bands / legendre_derivative_block (reference
 * legendre.py:76-146) so it can be checked against them. */
int arwen_sht_plan_table_f64(const arwen_sht_plan* plan, int kind, double* out);

/* Single operations (one grouped-GEMM launch each).  The plan keeps a
 * one-job batch per (operation, count, stream): the first call of a key
 * allocates its workspaces and uploads its work lists (and waits for that);
 * every later call rebinds the caller's buffers on the host and launches,
 * with no allocation and no synchronization.  At most 16 keys are kept per
 * plan; evicting one waits for its stream. */
/* forward (reference transform.py:1486-1490) */
int arwen_sht_analysis(arwen_sht_plan* plan, const float* grid, void* coeff, int count,
                       void* stream);
/* inverse (transform.py:1742-1746) */
int arwen_sht_synthesis(arwen_sht_plan* plan, const void* coeff, float* grid, int count,
                        void* stream);
/* inverse_zonal_derivative (transform.py:1752-1760) */
int arwen_sht_zonal_derivative(arwen_sht_plan* plan, const void* coeff, float* grid, int count,
                               void* stream);
/* inverse_meridional_derivative (transform.py:1748-1750) */
int arwen_sht_meridional_derivative(arwen_sht_plan* plan, const void* coeff, float* grid,
                                    int count, void* stream);
/* gradient (transform.py:1762-1768): east = d/dlambda / (a cos), north = d/dphi / a */
int arwen_sht_gradient(arwen_sht_plan* plan, const void* coeff, float* east, float* north,
                       int count, void* stream);
/* inverse with both gradients of the same field (ARWEN_SHT_OP_SYNTHESIS_GRADIENT):
 * grid, east = d/dlambda / (a cos), north = d/dphi / a, from two plane
 * contractions (reference transform.py:1742-1768, as the dycore calls
 * inverse and gradient on temperature and ln ps, dynamics.py:907-1047). */
int arwen_sht_synthesis_gradient(arwen_sht_plan* plan, const void* coeff, float* grid,
                                 float* east, float* north, int count, void* stream);
/* forward about offsets: forward(grid - grid_offset[f]) with coeff_offset[f]
 * added to the (0, 0) coefficient (either offset may be NULL; device arrays
 * of count float32 / complex64): the theta analysis offsets of the reference
 * (semilag/step.py:128-163, 610-621) and lnps_from_anomaly_coefficients
 * after an analysis (lnps_reference.py:47-49), fused into the fold and the
 * unpack. */
int arwen_sht_analysis_about(arwen_sht_plan* plan, const float* grid, void* coeff, int count,
                             const float* grid_offset, const void* coeff_offset, void* stream);
/* inverse of coeff with coeff_offset[f] (complex64, or NULL) added to its
 * (0, 0) coefficient in the pack (lnps_anomaly_coefficients before an
 * inverse, reference lnps_reference.py:32-44, dynamics.py:978-980); flags
 * ARWEN_SHT_JOB_EXP writes exp of the grid (ps = exp(ln ps),
 * dynamics.py:964-965). */
int arwen_sht_synthesis_about(arwen_sht_plan* plan, const void* coeff, float* grid, int count,
                              const void* coeff_offset, int flags, void* stream);
/* VorticityDivergenceOperator.divergence_from_fourier (reference
 * vector.py:136-166): divergence from (u, v) reference waists. */
int arwen_sht_divergence_from_waist(arwen_sht_plan* plan, const void* u_waist,
                                    const void* v_waist, void* div, int count, void* stream);
/* VorticityDivergenceOperator.wind_from_vordiv (reference vector.py:36-52) */
int arwen_sht_uv_from_vordiv(arwen_sht_plan* plan, const void* vort, const void* div, float* u,
                             float* v, int count, void* stream);
/* VorticityDivergenceOperator.vordiv_from_wind (reference vector.py:54-73, 113-134) */
int arwen_sht_vordiv_from_uv(arwen_sht_plan* plan, const float* u, const float* v, void* vort,
                             void* div, int count, void* stream);

/* The Fourier waist (count, nlat, T+1) complex64 of the reference's banded
 * pipeline.
 * contract_to_waist (reference transform.py:1629-1652): the Legendre
 * contraction of `coeff` with the basis (kind 0), with the basis after a
 * multiplication by i*m (kind 1; gradient_waists' east half, :1770-1787) or
 * with dP/dphi (kind 2; its north half). */
int arwen_sht_synthesis_to_waist(arwen_sht_plan* plan, const void* coeff, void* waist, int count,
                                 int kind, void* stream);
/* contract_waist / contract_values (transform.py:1435-1472): analysis of a
 * filled waist, m = 0 realified. */
int arwen_sht_analysis_from_waist(arwen_sht_plan* plan, const void* waist, void* coeff, int count,
                                  void* stream);
/* VorticityDivergenceOperator.vordiv_from_wind_waist / _vordiv_from_fourier
 * (reference vector.py:75-134): vorticity and divergence from the waists of
 * u and v, count levels each. */
int arwen_sht_vordiv_from_waist(arwen_sht_plan* plan, const void* u_waist, const void* v_waist,
                                void* vort, void* div, int count, void* stream);
/* FourierWaist.fill_band / fourier_waist (transform.py:483-525, 1294-1404):
 * `rows` (nlead, nrows, nlon) float32 -> waist rows row0 .. row0+nrows of
 * each lead index of `waist` (nlead, nlat, T+1): rfft / nlon, m <= T. */
int arwen_sht_waist_from_rows(arwen_sht_plan* plan, const float* rows, void* waist, int nlead,
                              int nrows, int row0, void* stream);
/* _rows_to_grid / waist_band_to_grid (transform.py:1654-1711): waist rows
 * row0 .. row0+nrows of each lead index -> `rows` (nlead, nrows, nlon)
 * float32: irfft of nlon x the retained orders. */
int arwen_sht_rows_from_waist(arwen_sht_plan* plan, const void* waist, float* rows, int nlead,
                              int nrows, int row0, void* stream);

/* Elementwise spectral operators on (count, T+1, T+1) complex64:
 * op 0 = project (transform.py:1124-1144), 1 = laplacian (:1798-1802),
 * 2 = inverse_laplacian (:1804-1811).  `in` may equal `out`. */
int arwen_sht_spectral_op(arwen_sht_plan* plan, int op, const void* in, void* out, int count,
                          void* stream);

/* Spectral fields at arbitrary points: sample_scalar, sample_gradient and
 * sample_wind (reference spectral/sampling.py:382-515).
 * kind 0: in0 (count, T+1, T+1) complex64 -> out0 (count, npoints) float64,
 *         the fields at the points (poles admitted);
 * kind 1: in0 -> out0 east, out1 north derivative per metre (points off the
 *         poles; east divides by radius * cos_lat);
 * kind 2: in0 vorticity, in1 divergence (count levels) -> out0 u, out1 v
 *         through psi, chi = inverse_laplacian (vorticity, divergence).
 * Points (device, float64, npoints each): sin_lat, cos_lat (kinds 1 and 2)
 * and lon_rad.  precision 0: float64 basis and sums; 1: float32 terms,
 * float64 sums over orders (the reference's "float32" sampler).
 * `workspace` (device) holds at least arwen_sht_sample_workspace_bytes
 * (the packed coefficients, and per-part sums of the order range). */
size_t arwen_sht_sample_workspace_bytes(const arwen_sht_plan* plan, int count, int kind,
                                        int npoints);
int arwen_sht_sample(arwen_sht_plan* plan, const void* in0, const void* in1, int count,
                     const double* sin_lat, const double* cos_lat, const double* lon_rad,
                     int npoints, int kind, int precision, void* workspace,
                     size_t workspace_bytes, double* out0, double* out1, void* stream);

/* arwen_sht_sample without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * workspace (arwen_sht_sample_workspace_bytes of the call) lives in a
 * library-owned array of ARWEN_SHT_SAMPLE_STATIC_BYTES per GPU, held until
 * the process ends; a call that needs more returns ARWEN_SHT_EINVAL without
 * launching anything.  The call allocates nothing.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams,
 * other threads) race.  Results are bit-identical to arwen_sht_sample's. */
#define ARWEN_SHT_SAMPLE_STATIC_BYTES (64u << 20)
int arwen_sht_sample_static(arwen_sht_plan* plan, const void* in0, const void* in1, int count,
                            const double* sin_lat, const double* cos_lat, const double* lon_rad,
                            int npoints, int kind, int precision, double* out0, double* out1,
                            void* stream);

/* The semi-implicit vertical-mode solves on spectral columns (reference
 * semi_implicit.py: VerticalModeSemiImplicit.solve_shifted :921-959 and
 * apply :887-919, each followed by _finish :815-843).  Inputs (device):
 * div, theta (nlev, T+1, T+1) and lnps (T+1, T+1) complex64; the vertical
 * operator's float32 matrices b_matrix (nlev, nlev+1), c_matrix (nlev+1,
 * nlev), m_matrix (nlev, nlev; read only when s_md != 0), the per-degree
 * Helmholtz inverses (T+1, nlev, nlev) and k2 (T+1).  Per column, with x =
 * [theta, lnps]:
 *   forcing = D + k2_n (s_bx B x + s_md M D);  d = inverse_n forcing;
 *   x' = x + s_c C (s_cd D + s_cn d);
 * then d(0,0) = D(0,0), x'(0,0) = x(0,0), and both projected.
 * solve_shifted: s_bx = s_c = tau w, s_md = s_cd = 0, s_cn = 1;
 * apply: s_bx = s_c = dt, s_md = -alpha (1 - alpha) dt^2, s_cd = 1 - alpha,
 * s_cn = alpha.  nlev <= 64.  Outputs must not alias inputs. */
int arwen_sht_vertical_solve(arwen_sht_plan* plan, int nlev, const void* div, const void* theta,
                             const void* lnps, const float* b_matrix, const float* c_matrix,
                             const float* m_matrix, const float* inverse, const float* k2,
                             float s_bx, float s_md, float s_c, float s_cd, float s_cn,
                             void* div_out, void* theta_out, void* lnps_out, void* stream);

/* linear_tendencies (semi_implicit.py:796-813): div_out = weight k2 B x,
 * (theta_out, lnps_out) = weight C D, each projected. */
int arwen_sht_vertical_tendencies(arwen_sht_plan* plan, int nlev, const void* div,
                                  const void* theta, const void* lnps, const float* b_matrix,
                                  const float* c_matrix, const float* k2, float weight,
                                  void* div_out, void* theta_out, void* lnps_out, void* stream);
/* half_apply (semi_implicit.py:845-878): the square root of the off-centred
 * Crank-Nicolson map from the operator's symmetric_half_maps, per degree
 * p_d, q_d (T+1, nlev, nlev) and p_x, q_x (T+1, nlev+1, nlev+1):
 *   d = P_D D + Q_D (k2 B x);  x' = P_x x + Q_x (C D);
 * degree zero pinned and projected as in _finish. */
int arwen_sht_vertical_half_apply(arwen_sht_plan* plan, int nlev, const void* div,
                                  const void* theta, const void* lnps, const float* b_matrix,
                                  const float* c_matrix, const float* p_d, const float* q_d,
                                  const float* p_x, const float* q_x, const float* k2,
                                  void* div_out, void* theta_out, void* lnps_out, void* stream);

/* A batch: an ordered list of jobs executed as
 *   pack -> ONE grouped synthesis GEMM -> inverse FFTs        (synthesis jobs)
 *   FFTs -> ONE grouped analysis GEMM -> unpack               (analysis jobs)
 * so every field, level, ensemble member and truncation of the batch rides one
 * persistent launch per direction.  Analysis jobs may read the grid outputs
 * of synthesis jobs of the same batch (a round trip).  Jobs may mix plans of
 * different truncations; all plans must share one precision.  No reference
 * counterpart: the reference steps fields in chunks and members serially
 * (reference da/ensemble.py:437-455). */
typedef struct arwen_sht_job {
  int op;                 /* ARWEN_SHT_OP_* */
  arwen_sht_plan* plan;
  int count;              /* fields (scalar ops) or levels (vector ops) */
  const void* in0;
  const void* in1;
  void* out0;
  void* out1;
  void* out2;             /* ARWEN_SHT_OP_SYNTHESIS_GRADIENT: north */
  int flags;              /* ARWEN_SHT_JOB_* */
  /* Offsets about which a field is transformed (NULL: none), device arrays
   * of `count` entries.  ANALYSIS: grid_offset[f] (float32) is subtracted
   * from field f's grid before its FFT, and coeff_offset[f] (complex64) is
   * added to its (0, 0) coefficient afterwards: forward(g - c) + c sqrt(4 pi)
   * with the theta analysis offsets of the reference (semilag/step.py:128-163,
   * 610-621).  SYNTHESIS, SYNTHESIS_GRADIENT: coeff_offset[f] is added to the
   * (0, 0) coefficient before the transform, as add_grid_constant does
   * (transform.py:1832-1835; lnps_reference.py:32-49). */
  const float* grid_offset;
  const void* coeff_offset;
} arwen_sht_job;

int arwen_sht_batch_create(const arwen_sht_job* jobs, int njobs, arwen_sht_batch** out);
int arwen_sht_batch_destroy(arwen_sht_batch* batch);
/* Enqueue the batch's kernels on `stream`. */
int arwen_sht_batch_run(arwen_sht_batch* batch, void* stream);
/* Capture the batch into a hipGraph (once); later runs launch the graph. */
int arwen_sht_batch_capture(arwen_sht_batch* batch, void* stream);
/* Launch the captured graph (captures first if needed). */
int arwen_sht_batch_launch(arwen_sht_batch* batch, void* stream);
/* Run the batch `reps` times with HIP events recorded on `stream` around each
 * phase and return the mean milliseconds per phase in ms[8]: pack, synthesis
 * GEMM, the vector jobs' FP64 low-order synthesis (its pack included),
 * inverse FFTs, forward FFTs, analysis GEMM, FP64 low-order analysis,
 * unpack.  Synchronizes on
 * the last event.  Measurement hook; no reference counterpart (the
 * reference's StepProfiler times whole sections, profile.py:87-330). */
int arwen_sht_batch_profile(arwen_sht_batch* batch, void* stream, int reps, double* ms);
/* Grouped-GEMM schedule of a batch, per direction (0 synthesis, 1 analysis):
 * column tile (32, 64 or 96 data columns per workgroup; 96 only with
 * split_acc = 0 in the split precisions) and persistent grid (workgroups,
 * 0 = 8 per CU); for the split precisions, whether the small products keep
 * their own accumulators (1, default) or share the main one (0).  Changing
 * it rebuilds the work lists and drops the captured graph.  No reference
 * counterpart (the reference has no kernel schedule). */
typedef struct arwen_sht_gemm_config {
  int cols[2];
  int grid[2];
  int split_acc;
} arwen_sht_gemm_config;
int arwen_sht_batch_get_config(const arwen_sht_batch* batch, arwen_sht_gemm_config* out);
int arwen_sht_batch_set_config(arwen_sht_batch* batch, const arwen_sht_gemm_config* config);
/* Autotuner: times each direction's GEMM launch `reps` times on `stream` for
 * every column tile and grid, keeps the fastest and returns it in `out`
 * (optional).  Synchronizes; overwrites the batch's intermediate buffers. */
int arwen_sht_batch_autotune(arwen_sht_batch* batch, void* stream, int reps,
                             arwen_sht_gemm_config* out);
/* Algorithmic work of the batch: Legendre FLOPs (parity folded, triangular,
 * all orders), FFT bytes (grid + waist once each) and the part of the
 * Legendre FLOPs that the FP64 low-order kernels do (the vector jobs' orders
 * m < fp64_orders; the grouped GEMMs do the rest) per run.  Any output
 * pointer may be NULL. */
int arwen_sht_batch_work(const arwen_sht_batch* batch, double* legendre_flops, double* fft_bytes,
                         double* fp64_low_flops);

const char* arwen_sht_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* ARWEN_SHT_H */
