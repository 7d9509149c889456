// C ABI of the spherical-harmonic transform engine (include/arwen_sht.h).
//
// A plan owns one truncation's device state: Gauss nodes and weights, the
// three Legendre tables in both GEMM layouts, the FFT twiddles and the
// per-row scale vectors.  A batch owns the workspaces, the grouped-GEMM work
// list of its jobs and, once captured, a hipGraph of its launches.
#include "arwen_sht.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "sht_kernels.h"
#include "sht_static.h"

namespace {

thread_local std::string g_error;

struct Failure {
  int code;
};

[[noreturn]] void fail(int code, const std::string& msg) {
  g_error = msg;
  throw Failure{code};
}

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    fail(e == hipErrorOutOfMemory ? ARWEN_SHT_ENOMEM : ARWEN_SHT_EHIP,
         std::string(what) + ": " + hipGetErrorString(e));
  }
}

template <class F>
int guarded(F&& f) {
  try {
    f();
    return ARWEN_SHT_OK;
  } catch (const Failure& x) {
    return x.code;
  } catch (const std::bad_alloc&) {
    g_error = "host allocation failed";
    return ARWEN_SHT_ENOMEM;
  } catch (...) {
    g_error = "unexpected C++ exception";
    return ARWEN_SHT_EHIP;
  }
}

// Set-up work of a single operation's first call goes to the caller's stream
// (zero fills and uploads ordered before its kernels, no device-wide wait);
// null elsewhere.  The host copies of what is uploaded there live in
// g_setup_keep until the owner takes them.
thread_local hipStream_t g_setup_stream = nullptr;
thread_local std::vector<std::vector<char>>* g_setup_keep = nullptr;

// Owning device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t n, bool zero = true) { alloc(n, zero); }
  void alloc(size_t n, bool zero = true) {
    release();
    bytes = n;
    if (n == 0) return;
    hip_check(hipMalloc(&p, n), "hipMalloc");
    if (zero) {
      if (g_setup_stream)
        hip_check(hipMemsetAsync(p, 0, n, g_setup_stream), "hipMemsetAsync");
      else
        hip_check(hipMemset(p, 0, n), "hipMemset");
    }
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

template <class T>
void upload(DevBuf& buf, const std::vector<T>& host) {
  buf.alloc(host.size() * sizeof(T), false);
  if (host.empty()) return;
  const size_t n = host.size() * sizeof(T);
  if (g_setup_stream && g_setup_keep) {
    // the copy may still be pending when this returns: copy from a kept buffer
    g_setup_keep->emplace_back(reinterpret_cast<const char*>(host.data()),
                               reinterpret_cast<const char*>(host.data()) + n);
    hip_check(hipMemcpyAsync(buf.p, g_setup_keep->back().data(), n, hipMemcpyHostToDevice,
                             g_setup_stream),
              "hipMemcpyAsync H2D");
  } else {
    hip_check(hipMemcpy(buf.p, host.data(), n, hipMemcpyHostToDevice), "hipMemcpy H2D");
  }
}

std::vector<int> factor_radices(int n) {
  std::vector<int> r;
  while (n % 4 == 0) { r.push_back(4); n /= 4; }
  while (n % 2 == 0) { r.push_back(2); n /= 2; }
  while (n % 3 == 0) { r.push_back(3); n /= 3; }
  for (int p = 5; n > 1; p += 2) {
    while (n % p == 0) { r.push_back(p); n /= p; }
    if (p * p > n && n > 1) { r.push_back(n); n = 1; }
  }
  return r;
}

}  // namespace

enum TableKind { kP = 0, kD = 1, kPm = 2 };

struct arwen_sht_batch;

struct arwen_sht_plan {
  int device = 0;
  int T = 0, nlat = 0, nlon = 0, JH = 0, JHP = 0, MB = 0;
  double radius = 0;
  int precision = 0;
  std::vector<int> koff;  // 2(T+1)+1 entries
  long long tab_elems = 0;
  int ldp = 0;            // packed floats per column
  int ldw = 0;            // waist floats per column
  std::vector<double> mu, w;
  DevBuf d_mu, d_w, d_koff, d_tw;
  DevBuf tab_syn[3], tab_ana[3];
  DevBuf scale_fold, scale_fold_vec, scale_fold_vec_cos, scale_one, scale_inv_a, scale_inv_acos;
  DevBuf scale_waist, scale_waist_vec;  // 2 pi w (x 1/a): folds of a reference waist (already / nlon)
  DevBuf sample_rec, sample_diag;       // point sampler recurrence tables, built on first use
  // vector operators' FP64 low orders m < m0: FP64 twiddles, per-row FP64
  // scales and the FP64 tables of those orders (kinds dP/dphi and P m / cos)
  int m0 = 0;
  DevBuf d_tw64;
  // FP64 row scales of the recurrence form (every order of u and v):
  // fold 2 pi w / (a nlon cos(phi)), unfold 1 / (a cos(phi))
  DevBuf scale_fold_vec64_cos, scale_inv_acos64;
  DevBuf low_syn[3], low_ana[3];
  sht::FftPlan fft{};
  int rp_fold = 4, rp_fold_vec = 4, rp_unfold = 4;  // folded rows (wavefronts) per FFT block
  int fft_gpb = 2;                                   // row groups walked by one FFT block
  // single-operation batches by (op, count, stream), most recent first
  std::list<std::pair<std::tuple<int, int, void*>, arwen_sht_batch*>> cache;
  std::mutex cache_mu;
  ~arwen_sht_plan();

  int kp(int m, int p) const { return koff[2 * m + p + 1] - koff[2 * m + p]; }
  size_t table_bytes() const {
    size_t per = precision == ARWEN_SHT_FP32 ? 4 : 6;
    size_t low = 0;
    for (int k = 0; k < 3; ++k) low += low_syn[k].bytes + low_ana[k].bytes;
    return 6 * (size_t)tab_elems * per + low;
  }
  const void* table(int kind, bool syn, long long* piece) const {
    *piece = tab_elems;
    return syn ? tab_syn[kind].p : tab_ana[kind].p;
  }
};

namespace {

void build_sample_tables(arwen_sht_plan* P);

void build_plan(arwen_sht_plan* P, const arwen_sht_plan_desc* d) {
  if (!d) fail(ARWEN_SHT_EINVAL, "plan descriptor is NULL");
  const int T = d->truncation;
  if (T < 1) fail(ARWEN_SHT_EINVAL, "truncation must be >= 1, got " + std::to_string(T));
  if (d->nlat < T + 1)
    fail(ARWEN_SHT_EINVAL, "nlat=" + std::to_string(d->nlat) + " is too small for T" +
                               std::to_string(T) + "; need at least " + std::to_string(T + 1));
  if (d->nlon % 2 || d->nlon < 2 * T + 2)
    fail(ARWEN_SHT_EINVAL, "nlon must be even and >= 2T+2, got " + std::to_string(d->nlon));
  if (!(d->radius_m > 0) || !std::isfinite(d->radius_m))
    fail(ARWEN_SHT_EINVAL, "radius_m must be finite and positive");
  if (d->precision != ARWEN_SHT_FP32 && d->precision != ARWEN_SHT_BF16X3 &&
      d->precision != ARWEN_SHT_BF16X6)
    fail(ARWEN_SHT_EINVAL, "precision must be ARWEN_SHT_FP32, _BF16X3 or _BF16X6");
  if ((d->sin_lat == nullptr) != (d->weights == nullptr))
    fail(ARWEN_SHT_EINVAL, "sin_lat and weights must be given together");
  hip_check(hipGetDevice(&P->device), "hipGetDevice");
  P->T = T;
  P->nlat = d->nlat;
  P->nlon = d->nlon;
  P->radius = d->radius_m;
  P->precision = d->precision;
  P->JH = (d->nlat + 1) / 2;
  P->JHP = SHT_ROUND_UP(P->JH, sht::kPad);
  P->MB = (T + 1 + sht::kMT - 1) / sht::kMT;
  // blocks hold degrees m..T+1: the vector operators' recurrence form
  // contracts the basis one degree past the truncation (the other operations
  // leave that row zero in their data and ignore it in their results)
  P->koff.assign(2 * (T + 1) + 1, 0);
  for (int m = 0; m <= T; ++m)
    for (int p = 0; p < 2; ++p)
      P->koff[2 * m + p + 1] = P->koff[2 * m + p] + sht::k_padded(T + 1, m, p);
  P->ldp = P->koff.back();
  P->tab_elems = (long long)P->ldp * P->JHP;
  P->ldw = P->MB * 2 * P->JHP * sht::kMT;

  // Nodes and weights.
  P->mu.resize(P->nlat);
  P->w.resize(P->nlat);
  P->d_mu.alloc(P->nlat * sizeof(double), false);
  P->d_w.alloc(P->nlat * sizeof(double), false);
  if (d->sin_lat) {
    for (int j = 0; j < P->nlat; ++j) {
      P->mu[j] = d->sin_lat[j];
      P->w[j] = d->weights[j];
    }
    for (int j = 0; j < P->nlat; ++j)
      if (P->mu[j] != -P->mu[P->nlat - 1 - j])
        fail(ARWEN_SHT_EINVAL, "sin_lat must be ascending and symmetric about the equator");
    hip_check(hipMemcpy(P->d_mu.p, P->mu.data(), P->nlat * 8, hipMemcpyHostToDevice), "copy mu");
    hip_check(hipMemcpy(P->d_w.p, P->w.data(), P->nlat * 8, hipMemcpyHostToDevice), "copy w");
  } else {
    hip_check(sht_launch_gauss_newton(P->nlat, P->d_mu.as<double>(), P->d_w.as<double>(), 0),
              "gauss_newton launch");
    hip_check(hipMemcpy(P->mu.data(), P->d_mu.p, P->nlat * 8, hipMemcpyDeviceToHost), "read mu");
    hip_check(hipMemcpy(P->w.data(), P->d_w.p, P->nlat * 8, hipMemcpyDeviceToHost), "read w");
  }
  upload(P->d_koff, P->koff);

  // FP64 vector orders (resolved below once the FFT path is known).
  if (d->fp64_orders < -1)
    fail(ARWEN_SHT_EINVAL, "fp64_orders must be -1 (default), 0 or positive, got " +
                               std::to_string(d->fp64_orders));
  {
    const int L = sht_fft_wave_L(P->nlon);
    const bool wave = L && sht_fold64_lds_bytes(P->nlon) <= 160 * 1024;
    int m0 = d->fp64_orders;
    if (m0 < 0) m0 = P->precision == ARWEN_SHT_FP32 ? 0 : sht::kMT * ((T + 1 + 191) / 192);
    m0 = SHT_ROUND_UP(m0, sht::kMT);
    P->m0 = wave ? std::min(m0, P->MB * sht::kMT) : 0;
  }

  // Legendre tables.
  sht::TableOut out{};
  out.elems = P->tab_elems;
  const bool fp32 = P->precision == ARWEN_SHT_FP32;
  for (int k = 0; k < 3; ++k) {
    size_t bytes = (size_t)P->tab_elems * (fp32 ? 4 : 6);
    P->tab_syn[k].alloc(bytes, true);
    P->tab_ana[k].alloc(bytes, true);
    if (fp32) {
      out.f32_syn[k] = P->tab_syn[k].as<float>();
      out.f32_ana[k] = P->tab_ana[k].as<float>();
    } else {
      out.bf_syn[k] = P->tab_syn[k].as<unsigned short>();
      out.bf_ana[k] = P->tab_ana[k].as<unsigned short>();
    }
  }
  if (P->m0 > 0) {
    // the basis over degrees m..T+1 in both layouts (the recurrence form in FP64)
    const size_t low_bytes = (size_t)P->m0 * (T + 2) * P->JHP * sizeof(double);
    P->low_syn[kP].alloc(low_bytes, true);
    P->low_ana[kP].alloc(low_bytes, true);
    out.f64_low_syn[kP] = P->low_syn[kP].as<double>();
    out.f64_low_ana[kP] = P->low_ana[kP].as<double>();
    out.low_m0 = P->m0;
  }
  hip_check(sht_launch_legendre_tables(T, P->nlat, P->JH, P->JHP, P->d_mu.as<double>(),
                                       P->d_koff.as<int>(), &out, 0),
            "legendre_tables launch");

  // FFT plan.
  std::vector<int> rad = factor_radices(P->nlon);
  if ((int)rad.size() > sht::kMaxStages) fail(ARWEN_SHT_EINVAL, "nlon has too many factors");
  P->fft.N = P->nlon;
  P->fft.nstages = (int)rad.size();
  for (size_t s = 0; s < rad.size(); ++s) P->fft.radix[s] = rad[s];
  std::vector<float2> tw(P->nlon);
  for (int k = 0; k < P->nlon; ++k) {
    double ang = -2.0 * M_PI * (double)k / (double)P->nlon;
    tw[k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
  }
  upload(P->d_tw, tw);
  P->fft.tw = P->d_tw.as<float2>();
  if (P->m0 > 0) {
    std::vector<double2> tw64(P->nlon);
    for (int k = 0; k < P->nlon; ++k) {
      const double ang = -2.0 * M_PI * (double)k / (double)P->nlon;
      tw64[k] = make_double2(std::cos(ang), std::sin(ang));
    }
    upload(P->d_tw64, tw64);
    P->fft.tw64 = P->d_tw64.as<double2>();
  }
  // Wavefronts (folded rows) per FFT block: as many as fit 80 KiB, so two
  // blocks share a CU; one row per block up to the device's whole LDS where
  // a single row needs more (nlon = 3072: 96 KiB for a u,v row pair).
  {
    int lds_max = 65536;
    (void)hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, P->device);
    const size_t budget = std::min<size_t>((size_t)lds_max, 81920);
    size_t need = budget;
    auto pick = [&](auto bytes) {
      int rp = 4;
      while (rp > 1 && bytes(rp) > budget) rp /= 2;
      if (bytes(rp) > (size_t)lds_max)
        fail(ARWEN_SHT_EINVAL, "nlon=" + std::to_string(P->nlon) + " needs more LDS than the device has");
      need = std::max(need, bytes(rp));
      return rp;
    };
    P->rp_fold = pick([&](int rp) { return sht_fold_lds_bytes(P->nlon, T, 0, rp); });
    P->rp_fold_vec = pick([&](int rp) { return sht_fold_lds_bytes(P->nlon, T, 1, rp); });
    P->rp_unfold = pick([&](int rp) { return sht_unfold_lds_bytes(P->nlon, T, rp); });
    hip_check(sht_fft_allow_lds(need, (size_t)lds_max), "FFT LDS attribute");
    if (P->m0 > 0) hip_check(sht_low_allow_lds((size_t)lds_max), "low-order LDS attribute");
    // wave-resident FFTs (N = 64 L) when their LDS fits the device
    const int L = sht_fft_wave_L(P->nlon);
    P->fft.wave_L = (L && sht_fft_wave_lds_bytes(P->nlon, 1) <= (size_t)lds_max) ? L : 0;
  }

  // Per folded-row scales.
  std::vector<float> sf(P->JHP, 0.f), sfv(P->JHP, 0.f), one(P->JHP, 1.f), inva(P->JHP, 0.f);
  std::vector<float> invacos(P->JHP, 0.f), sfvc(P->JHP, 0.f);
  std::vector<float> sw(P->JHP, 0.f), swv(P->JHP, 0.f);
  for (int i = 0; i < P->JH; ++i) {
    const int jn = P->nlat - 1 - i;
    double wq = P->w[jn];
    if (jn == i) wq *= 0.5;  // the equator row is its own mirror
    sf[i] = (float)(2.0 * M_PI * wq / P->nlon);
    sfv[i] = (float)(2.0 * M_PI * wq / (P->radius * P->nlon));
    sw[i] = (float)(2.0 * M_PI * wq);
    swv[i] = (float)(2.0 * M_PI * wq / P->radius);
    inva[i] = (float)(1.0 / P->radius);
    const double mu = P->mu[jn];
    invacos[i] = (float)(1.0 / (P->radius * std::sqrt(std::max(0.0, 1.0 - mu * mu))));
    sfvc[i] = (float)(2.0 * M_PI * wq / (P->radius * P->nlon) /
                      std::sqrt(std::max(0.0, 1.0 - mu * mu)));
  }
  upload(P->scale_fold, sf);
  upload(P->scale_fold_vec, sfv);
  upload(P->scale_fold_vec_cos, sfvc);
  upload(P->scale_waist, sw);
  upload(P->scale_waist_vec, swv);
  upload(P->scale_one, one);
  upload(P->scale_inv_a, inva);
  upload(P->scale_inv_acos, invacos);
  if (P->m0 > 0) {

    std::vector<double> sfvc64(P->JHP, 0.0), invac64(P->JHP, 0.0);
    for (int i = 0; i < P->JH; ++i) {
      const int jn = P->nlat - 1 - i;
      double wq = P->w[jn];
      if (jn == i) wq *= 0.5;
      const double mu = P->mu[jn];
      const double cosl = std::sqrt(std::max(0.0, 1.0 - mu * mu));
      sfvc64[i] = 2.0 * M_PI * wq / (P->radius * P->nlon) / cosl;
      invac64[i] = 1.0 / (P->radius * cosl);
    }
    upload(P->scale_fold_vec64_cos, sfvc64);
    upload(P->scale_inv_acos64, invac64);
  }
  // the point sampler's recurrence coefficients, so that sampling allocates nothing
  build_sample_tables(P);
  // this device's copy of arwen_sht_sample_static's array, resolved once
  hip_check(sht_static_setup(P->device), "static sampler workspace");
  hip_check(hipDeviceSynchronize(), "plan build");
}

}  // namespace

// ---------------------------------------------------------------------------
// Batches
// ---------------------------------------------------------------------------

// One work group of the grouped GEMM before it is cut into panels: up to kMT
// consecutive orders of one job and one output parity, with one or two
// segments (table kind, data) accumulated into the same output.
struct HostGroup {
  int nseg = 0, N = 0, mcount = 0;
  const char* A[2][sht::kMT] = {};  // table block of (seg, order), tiled layout
  int K[2][sht::kMT] = {};          // k extent: SYN degrees KP(m, p); ANA JHP
  int Arows[2][sht::kMT] = {};      // rows of the table block: SYN JHP; ANA KP(m, q)
  const char* B[2][sht::kMT] = {};  // SYN packed block of (seg, order); ANA waist block of seg
  float* C[sht::kMT] = {};          // SYN C[0] = waist block; ANA packed block of the order
  int R[sht::kMT] = {};             // output rows per order (SYN JHP; ANA KP(m, q))
  int cmap[2] = {0, 0};             // B column map per segment (GemmStep::cmap)
  int bcs = 0, ccs = 0;             // column strides (floats)
  long long piece_bytes = 0;        // table piece stride (split modes)
  size_t esz = 4;                   // table element bytes (4 fp32, 2 bf16 pieces); data are fp32
};

// Device work lists of one GEMM direction, built for one column tile.
struct GemmLists {
  DevBuf panels, steps, tiles;
  int ntiles = 0;
  int bn = 0;
};

struct JobState {
  arwen_sht_job job;
  DevBuf packed_a;            // synthesis input (packed coefficients)
  DevBuf waist_a, waist_b;    // synthesis output (b: gradient north) / analysis input
  DevBuf packed_out;          // analysis output
  DevBuf low;                 // vector jobs: FP64 low waist [level][u|v][m < m0][p][i]
  DevBuf low_ab;              // vordiv_from_uv: FP64 A, B [level][A|B][n < T+2][m < m0]
  int ncols = 0;
};

struct arwen_sht_batch {
  int device = 0;
  int precision = 0;
  std::vector<std::vector<char>> setup_keep;  // host sources of stream-ordered uploads
  std::vector<std::unique_ptr<JobState>> jobs;
  std::vector<sht::PackArgs> packs;
  std::vector<sht::UnfoldArgs> unfolds;
  std::vector<sht::FoldArgs> folds;
  std::vector<sht::UnpackArgs> unpacks;
  std::vector<sht::WaistUnfoldArgs> waist_unfolds;  // synthesis into a reference waist
  std::vector<sht::WaistFoldArgs> waist_folds;      // analysis from a reference waist
  std::vector<sht::LowSynArgs> low_syns;            // vector jobs' FP64 low orders
  std::vector<sht::LowAnaArgs> low_anas;
  std::vector<HostGroup> groups[2];  // GEMM work groups per direction (0 SYN, 1 ANA)
  GemmLists lists[2];
  arwen_sht_gemm_config config{{32, 32}, {0, 0}, 1};
  double legendre_flops = 0, fft_bytes = 0;
  double low_flops = 0;  // the part of legendre_flops the FP64 low-order kernels do
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~arwen_sht_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

arwen_sht_plan::~arwen_sht_plan() {
  for (auto& e : cache) delete e.second;
}

namespace {

bool is_synthesis(int op) {
  return (op >= ARWEN_SHT_OP_SYNTHESIS && op <= ARWEN_SHT_OP_UV_FROM_VORDIV) ||
         (op >= ARWEN_SHT_OP_SYNTHESIS_TO_WAIST && op <= ARWEN_SHT_OP_MERIDIONAL_TO_WAIST) ||
         op == ARWEN_SHT_OP_SYNTHESIS_GRADIENT;
}
bool to_waist(int op) {
  return op >= ARWEN_SHT_OP_SYNTHESIS_TO_WAIST && op <= ARWEN_SHT_OP_MERIDIONAL_TO_WAIST;
}
bool from_waist(int op) {
  return op == ARWEN_SHT_OP_ANALYSIS_FROM_WAIST || op == ARWEN_SHT_OP_VORDIV_FROM_WAIST ||
         op == ARWEN_SHT_OP_DIVERGENCE_FROM_WAIST;
}

constexpr int kXcds = 8;  // MI355X: 8 XCDs x 32 CUs

// Cuts the work groups of one direction into panels (group x 64-row tile),
// their k-step lists and (panel, column tile) work items, in XCD-aware order.
// Workgroups are dispatched round-robin over the 8 XCDs and the kernels walk
// the list with a grid stride that is a multiple of 8, so list position t
// always runs on XCD t % 8.  Every work item of one group (same orders, same
// tables, same data columns) goes to one XCD, where its tiles run side by
// side and share that XCD's L2; groups are balanced over the XCDs longest
// first.  Short lists are padded with idle items (-1).
void build_lists(int dir, const std::vector<HostGroup>& groups, int bn, GemmLists& out) {
  const int TR = sht::kGemmBM;
  std::vector<sht::GemmPanel> panels;
  std::vector<sht::GemmStep> steps;
  std::vector<long long> gwork(groups.size(), 0);
  std::vector<std::vector<int2>> per_group(groups.size());
  for (size_t gid = 0; gid < groups.size(); ++gid) {
    const HostGroup& g = groups[gid];
    int Rmax = 0;
    for (int z = 0; z < g.mcount; ++z) Rmax = std::max(Rmax, g.R[z]);
    for (int r0 = 0; r0 < Rmax; r0 += TR) {
      sht::GemmPanel pn{};
      pn.step0 = (int)steps.size();
      pn.N = g.N;
      pn.bcs = g.bcs;
      pn.ccs = g.ccs;
      pn.piece_bytes = g.piece_bytes;
      for (int z = 0; z < sht::kMT; ++z) {
        pn.rows[z] = z < g.mcount ? std::max(0, std::min(TR, g.R[z] - r0)) : 0;
        if (dir == 0)
          pn.c[z] = z == 0 ? (unsigned long long)(g.C[0] + (size_t)r0 * sht::kMT) : 0ull;
        else
          pn.c[z] = z < g.mcount ? (unsigned long long)(g.C[z] + r0) : 0ull;
      }
      for (int sg = 0; sg < g.nseg; ++sg) {
        int kmax = 0;
        for (int z = 0; z < g.mcount; ++z) kmax = std::max(kmax, g.K[sg][z]);
        // every k-chunk is a trip of kMT steps, step z for order z (the
        // kernel fixes each step's order at compile time); an order that
        // does not reach this chunk or row tile gets a null step (extent 0:
        // its loads return zeros).  ANA: the chunk's B tile, shared by the
        // orders, comes with step 0.
        for (int k0 = 0; k0 < kmax; k0 += 32) {
          for (int z = 0; z < sht::kMT; ++z) {
            sht::GemmStep st{};
            st.z = z;
            st.cmap = g.cmap[sg];
            const bool live = z < g.mcount && k0 < g.K[sg][z] &&
                              (dir == 0 || g.Arows[sg][z] > r0);
            if (live) {
              const int chunk_rows = g.Arows[sg][z];
              st.a = (unsigned long long)(g.A[sg][z] +
                                          ((size_t)k0 * chunk_rows + (size_t)r0 * 32) * g.esz);
              st.a_bytes = (unsigned)(std::min(TR, chunk_rows - r0) * 32 * g.esz);
            } else {
              st.a = (unsigned long long)g.A[sg][0];
            }
            if (dir == 0) {
              st.b = (unsigned long long)(live ? g.B[sg][z] + (size_t)k0 * sizeof(float)
                                               : g.B[sg][0]);
              st.b_klen = live ? g.K[sg][z] - k0 : 0;
            } else if (z == 0) {
              st.b = (unsigned long long)(g.B[sg][0] + (size_t)k0 * sht::kMT * sizeof(float));
              st.b_klen = g.K[sg][0] - k0;
            }
            steps.push_back(st);
          }
        }
      }
      pn.nsteps = (int)steps.size() - pn.step0;
      bool any = false;
      for (int e = pn.step0; e < (int)steps.size(); ++e) any = any || steps[e].a_bytes > 0;
      if (!any) {  // nothing reaches this row tile
        steps.resize(pn.step0);
        continue;
      }
      if (pn.nsteps + 2 > sht::kGemmMaxSteps)
        fail(ARWEN_SHT_EINVAL, "truncation too large for the GEMM step list (" +
                                   std::to_string(pn.nsteps) + " steps in one panel, at most " +
                                   std::to_string(sht::kGemmMaxSteps - 2) + ")");
      // two null steps: the kernel's loads run ahead of the list's end
      sht::GemmStep null{};
      null.a = steps[pn.step0].a;
      null.b = steps[pn.step0].b ? steps[pn.step0].b : steps[pn.step0].a;
      steps.push_back(null);
      steps.push_back(null);
      const int pid = (int)panels.size();
      panels.push_back(pn);
      for (int c0 = 0; c0 < g.N; c0 += bn) {
        gwork[gid] += (long long)pn.nsteps * std::min(bn, g.N - c0);
        per_group[gid].push_back(int2{pid, c0});
      }
    }
  }
  const int nx = kXcds;
  std::vector<int> order(groups.size());
  for (size_t g = 0; g < order.size(); ++g) order[g] = (int)g;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return gwork[a] > gwork[b]; });
  std::vector<long long> load(nx, 0);
  std::vector<std::vector<int2>> lists(nx);
  for (int g : order) {
    if (per_group[g].empty()) continue;
    const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[x] += gwork[g];
    for (auto& tt : per_group[g]) lists[x].push_back(tt);
  }
  size_t longest = 0;
  for (auto& l : lists) longest = std::max(longest, l.size());
  std::vector<int2> t;
  t.reserve(longest * nx);
  for (size_t j = 0; j < longest; ++j)
    for (int x = 0; x < nx; ++x) t.push_back(j < lists[x].size() ? lists[x][j] : int2{-1, 0});
  upload(out.panels, panels);
  upload(out.steps, steps);
  upload(out.tiles, t);
  out.ntiles = (int)t.size();
  out.bn = bn;
}

const char* table_block(const arwen_sht_plan* P, int kind, bool syn, int m, int p,
                        long long* piece_bytes, size_t* esz) {
  long long piece;
  const char* base = static_cast<const char*>(P->table(kind, syn, &piece));
  *esz = P->precision == ARWEN_SHT_FP32 ? 4 : 2;
  *piece_bytes = piece * (long long)*esz;
  return base + (size_t)P->koff[2 * m + p] * P->JHP * *esz;
}

// Element offset of the waist block (m-tile of m, parity p) within a column,
// layout W[c][m/kMT][p][i][m%kMT].
size_t waist_offset(const arwen_sht_plan* P, int m, int p) {
  return ((size_t)((m / sht::kMT) * 2 + p) * P->JHP) * sht::kMT;
}

// Column maps of the vector operators' second data arrangement (see
// GemmStep::cmap): synthesis [i chi, i psi] = [-a3, a2, a1, -a0] of the packed
// [-psi, chi]; analysis [U, -V] = [a3, -a2, -a1, a0] of the waist [iV, iU].
constexpr int kMapSynVec = 1 | (1 << 1) | (1 << 4);
constexpr int kMapAnaVec = 1 | (1 << 2) | (1 << 3);

// GEMM data operand (fp32 in every precision mode; the split modes cut it
// into bf16 pieces as it enters the GEMM's LDS tiles).
struct DataRef {
  const char* base;
};

// Synthesis group: segments (table kind, parity, packed data, column map).
struct SynSeg {
  int kind;
  int pflip;  // data/table parity = q ^ pflip
  DataRef packed;
  int cmap;
};

void add_syn_group(std::vector<HostGroup>& b, const arwen_sht_plan* P, int mb, int q, int ncols,
                   float* waist, const SynSeg* segs, int nseg) {
  HostGroup g;
  g.nseg = nseg;
  g.N = ncols;
  g.mcount = std::min(sht::kMT, P->T + 1 - mb * sht::kMT);
  g.ccs = P->ldw;
  g.bcs = P->ldp;
  g.C[0] = waist + waist_offset(P, mb * sht::kMT, q);
  for (int sg = 0; sg < nseg; ++sg) g.cmap[sg] = segs[sg].cmap;
  for (int mm = 0; mm < g.mcount; ++mm) {
    const int m = mb * sht::kMT + mm;
    g.R[mm] = P->JHP;
    for (int sg = 0; sg < nseg; ++sg) {
      const int p = q ^ segs[sg].pflip;
      g.A[sg][mm] = table_block(P, segs[sg].kind, true, m, p, &g.piece_bytes, &g.esz);
      g.Arows[sg][mm] = P->JHP;
      g.B[sg][mm] = segs[sg].packed.base + (size_t)P->koff[2 * m + p] * sizeof(float);
      g.K[sg][mm] = P->kp(m, p);
    }
  }
  b.push_back(g);
}

// Analysis group: segments (table kind, waist parity flip, waist, column map).
struct AnaSeg {
  int kind;
  int pflip;
  DataRef waist;
  int cmap;
};

void add_ana_group(std::vector<HostGroup>& b, const arwen_sht_plan* P, int mb, int q, int ncols,
                   float* packed_out, const AnaSeg* segs, int nseg) {
  HostGroup g;
  g.nseg = nseg;
  g.N = ncols;
  g.mcount = std::min(sht::kMT, P->T + 1 - mb * sht::kMT);
  g.ccs = P->ldp;
  g.bcs = P->ldw;
  for (int sg = 0; sg < nseg; ++sg) {
    g.B[sg][0] = segs[sg].waist.base +
                 waist_offset(P, mb * sht::kMT, q ^ segs[sg].pflip) * sizeof(float);
    g.cmap[sg] = segs[sg].cmap;
  }
  for (int mm = 0; mm < g.mcount; ++mm) {
    const int m = mb * sht::kMT + mm;
    g.R[mm] = P->kp(m, q);
    g.C[mm] = packed_out + P->koff[2 * m + q];
    for (int sg = 0; sg < nseg; ++sg) {
      g.A[sg][mm] = table_block(P, segs[sg].kind, false, m, q, &g.piece_bytes, &g.esz);
      g.Arows[sg][mm] = P->kp(m, q);
      g.K[sg][mm] = P->JHP;
    }
  }
  b.push_back(g);
}

// One job's share of a batch.  prepare: allocate its workspaces, add its
// grouped-GEMM work groups and its algorithmic work (once, at batch
// creation); always: append its launch arguments, which carry the caller's
// buffers (so a single-operation batch is rebound to new buffers without
// allocating).  The work groups reference only workspaces and tables.
void setup_job(arwen_sht_batch* B, JobState* js, bool prepare) {
  auto& syn = B->groups[0];
  auto& ana = B->groups[1];
  const arwen_sht_job& J = js->job;
  const arwen_sht_plan* P = J.plan;
  const int n = J.count;
  const int T = P->T;
  const bool vec = J.op == ARWEN_SHT_OP_UV_FROM_VORDIV || J.op == ARWEN_SHT_OP_VORDIV_FROM_UV ||
                   J.op == ARWEN_SHT_OP_VORDIV_FROM_WAIST;
  const bool sgrad = J.op == ARWEN_SHT_OP_SYNTHESIS_GRADIENT;
  const double pcf = 4.0 * (P->nlat / 2.0) * (T + 1.0) * (T + 2.0) / 2.0;  // FLOP per PC
  // the orders m < m0 of that (vector jobs in the split modes: FP64 kernels)
  const double pcf_low =
      4.0 * (P->nlat / 2.0) * (P->m0 * (T + 1.0) - P->m0 * (P->m0 - 1.0) / 2.0);
  // the first order tile's share (orders m < 4), four contractions per level
  // there in the plain FP32 recurrence form
  const double pcf_t0 = 4.0 * (P->nlat / 2.0) * (sht::kMT * (T + 1.0) - sht::kMT * (sht::kMT - 1.0) / 2.0);
  const double fftb = 4.0 * P->nlat * P->nlon + 8.0 * P->nlat * (T + 1.0);
  if (n == 0) return;
  if (is_synthesis(J.op)) {
    // ---- pack ----
    const int ncol = vec ? 4 : 2;
    js->ncols = ncol * n;
    if (prepare) js->packed_a.alloc((size_t)js->ncols * P->ldp * 4);
    if (prepare) js->waist_a.alloc((size_t)js->ncols * P->ldw * 4);
    const bool grad = J.op == ARWEN_SHT_OP_GRADIENT;
    if ((grad || sgrad) && prepare) js->waist_b.alloc((size_t)js->ncols * P->ldw * 4);
    sht::PackArgs pa{};
    pa.koff = P->d_koff.as<int>();
    pa.T = T;
    pa.mtiles = (T + 1 + 31) / 32;
    pa.job.in0 = static_cast<const float2*>(J.in0);
    pa.job.in1 = static_cast<const float2*>(J.in1);
    pa.job.out0 = js->packed_a.as<float>();
    pa.job.nfields = n;
    pa.job.col0 = 0;
    pa.job.ldp = P->ldp;
    pa.job.radius = (float)P->radius;
    // u, v with FP64 low orders: the orders m >= m0 in the recurrence form;
    // without them (plain FP32) on a wave-resident grid: the recurrence form
    // from m = 4 on, the first order tile in the two-segment form
    const bool low = J.op == ARWEN_SHT_OP_UV_FROM_VORDIV && P->m0 > 0;
    const bool recp = J.op == ARWEN_SHT_OP_UV_FROM_VORDIV && P->m0 == 0 && P->fft.wave_L > 0;
    pa.job.mode = (J.op == ARWEN_SHT_OP_ZONAL_DERIVATIVE || J.op == ARWEN_SHT_OP_ZONAL_TO_WAIST)
                      ? sht::kIM
                  : low || recp                      ? sht::kVectorRec
                  : vec                              ? sht::kVector
                  : (J.flags & ARWEN_SHT_JOB_INVLAP) ? sht::kInvLap
                                                     : sht::kPlain;
    pa.job.off00 = static_cast<const float2*>(J.coeff_offset);
    pa.job.m0 = low ? P->m0 : 0;
    pa.job.mdp = recp ? sht::kMT : 0;
    B->packs.push_back(pa);
    // ---- FP64 low orders of u, v (orders m < m0) ----
    if (low) {
      if (prepare) js->low.alloc((size_t)n * 2 * P->m0 * 2 * P->JHP * sizeof(double2));
      sht::LowSynArgs ls{};
      ls.vort = static_cast<const float2*>(J.in0);
      ls.div = static_cast<const float2*>(J.in1);
      ls.p = P->low_syn[kP].as<double>();
      ls.low = js->low.as<double2>();
      ls.radius = P->radius;
      ls.T = T;
      ls.JH = P->JH;
      ls.JHP = P->JHP;
      ls.m0 = P->m0;
      ls.nlev = n;
      B->low_syns.push_back(ls);
    }
    // ---- grouped GEMM: one group per (m-tile, output parity) ----
    const DataRef pk{js->packed_a.as<char>()};
    for (int mb = low ? P->m0 / sht::kMT : 0; prepare && mb < P->MB; ++mb)
      for (int q = 0; q < 2; ++q) {
        if (low || (recp && mb >= 1)) {
          // recurrence form: a cos(phi) [u, v] = P [U, V] (pack_vector_rec_kernel)
          SynSeg sg[1] = {{kP, 0, pk, 0}};
          add_syn_group(syn, P, mb, q, js->ncols, js->waist_a.as<float>(), sg, 1);
        } else if (J.op == ARWEN_SHT_OP_UV_FROM_VORDIV) {
          // U,V = -dP/dphi (psi) + i P m/cos (chi), i P m/cos (psi) + dP/dphi (chi):
          // dP/dphi flips the parity, so its segment reads the other parity.
          SynSeg sg[2] = {{kD, 1, pk, 0}, {kPm, 0, pk, kMapSynVec}};
          add_syn_group(syn, P, mb, q, js->ncols, js->waist_a.as<float>(), sg, 2);
        } else if (J.op == ARWEN_SHT_OP_MERIDIONAL_DERIVATIVE ||
                   J.op == ARWEN_SHT_OP_MERIDIONAL_TO_WAIST) {
          SynSeg sg[1] = {{kD, 1, pk, 0}};
          add_syn_group(syn, P, mb, q, js->ncols, js->waist_a.as<float>(), sg, 1);
        } else if (grad) {
          SynSeg se[1] = {{kPm, 0, pk, 0}};
          add_syn_group(syn, P, mb, q, js->ncols, js->waist_a.as<float>(), se, 1);
          SynSeg sn[1] = {{kD, 1, pk, 0}};
          add_syn_group(syn, P, mb, q, js->ncols, js->waist_b.as<float>(), sn, 1);
        } else if (sgrad) {  // the basis waist feeds the grid and (times i m) the east
          SynSeg sp[1] = {{kP, 0, pk, 0}};
          add_syn_group(syn, P, mb, q, js->ncols, js->waist_a.as<float>(), sp, 1);
          SynSeg sn[1] = {{kD, 1, pk, 0}};
          add_syn_group(syn, P, mb, q, js->ncols, js->waist_b.as<float>(), sn, 1);
        } else {  // SYNTHESIS, ZONAL_DERIVATIVE and their waist forms
          SynSeg sg[1] = {{kP, 0, pk, 0}};
          add_syn_group(syn, P, mb, q, js->ncols, js->waist_a.as<float>(), sg, 1);
        }
      }
    // algorithmic FLOPs: one plane contraction per table application
    {
      double pcs = (J.op == ARWEN_SHT_OP_UV_FROM_VORDIV) ? 4.0 * n
                   : (grad || sgrad)                     ? 2.0 * n
                                                         : 1.0 * n;
      if (prepare && low) {  // two per level: the GEMM above m0, the FP64 kernel below
        B->legendre_flops += 2.0 * n * pcf;
        B->low_flops += 2.0 * n * pcf_low;
      } else if (prepare && recp) {  // four per level in the first order tile, two above
        B->legendre_flops += 2.0 * n * pcf + 2.0 * n * pcf_t0;
      } else if (prepare) {
        B->legendre_flops += pcs * pcf;
      }
    }
    // ---- inverse FFTs ----
    auto unfold = [&](const DevBuf& waist, void* grid, int col0, int fstride, const DevBuf& sc,
                      int imul, int post_exp = 0) {
      sht::UnfoldArgs ua{};
      ua.fft = P->fft;
      ua.T = T;
      ua.nlat = P->nlat;
      ua.JH = P->JH;
      ua.JHP = P->JHP;
      ua.rp = P->rp_unfold;
      ua.ngroups = (P->JH + ua.rp - 1) / ua.rp;
      ua.gpb = P->fft_gpb;
      ua.nchunks = (ua.ngroups + ua.gpb - 1) / ua.gpb;
      ua.job.waist = waist.as<float>();
      ua.job.grid = static_cast<float*>(grid);
      ua.job.rowscale = sc.as<float>();
      ua.job.nfields = n;
      ua.job.col0 = col0;
      ua.job.fstride = fstride;
      ua.job.ldw = P->ldw;
      ua.job.imul = imul;
      ua.job.post_exp = post_exp;
      if (low) {  // FP64 unfold; orders m < m0 from the low waist of u (col0 0) or v
        const int f = col0 == 0 ? 0 : 1;
        ua.job.f64 = 1;
        ua.job.low = js->low.as<double2>() + (size_t)f * P->m0 * 2 * P->JHP;
        ua.job.low_stride = 2 * P->m0 * 2 * P->JHP;
        // both parts in the recurrence form: a cos(phi) (u, v)
        ua.job.rowscale64 = P->scale_inv_acos64.as<double>();
        ua.job.rowscale64_hi = P->scale_inv_acos64.as<double>();
        ua.job.m0 = P->m0;
      }
      B->unfolds.push_back(ua);
      if (prepare) B->fft_bytes += n * fftb;
    };
    if (to_waist(J.op)) {
      // the reference's waist (*lead, nlat, T+1): north = E + O, south = E - O
      sht::WaistUnfoldArgs wu{};
      wu.waist = js->waist_a.as<float>();
      wu.out = static_cast<float2*>(J.out0);
      wu.T = T;
      wu.nlat = P->nlat;
      wu.JH = P->JH;
      wu.JHP = P->JHP;
      wu.ldw = P->ldw;
      wu.nfields = n;
      wu.col0 = 0;
      wu.fstride = 2;
      B->waist_unfolds.push_back(wu);
    } else if (J.op == ARWEN_SHT_OP_UV_FROM_VORDIV) {
      unfold(js->waist_a, J.out0, 0, 4, P->scale_inv_a, 0);
      unfold(js->waist_a, J.out1, 2, 4, P->scale_inv_a, 0);
      if (recp)  // the recurrence form's orders: a cos(phi) (u, v)
        for (size_t k = B->unfolds.size() - 2; k < B->unfolds.size(); ++k) {
          B->unfolds[k].job.rowscale_hi = P->scale_inv_acos.as<float>();
          B->unfolds[k].job.m_hi = sht::kMT;
        }
    } else if (grad) {
      unfold(js->waist_a, J.out0, 0, 2, P->scale_inv_a, 1);
      unfold(js->waist_b, J.out1, 0, 2, P->scale_inv_a, 0);
    } else if (sgrad) {
      unfold(js->waist_a, J.out0, 0, 2, P->scale_one, 0);
      unfold(js->waist_a, J.out1, 0, 2, P->scale_inv_acos, 2);  // i m / (a cos)
      unfold(js->waist_b, J.out2, 0, 2, P->scale_inv_a, 0);
    } else {
      unfold(js->waist_a, J.out0, 0, 2, P->scale_one, 0, (J.flags & ARWEN_SHT_JOB_EXP) ? 1 : 0);
    }
  } else {
    // ---- FFT + fold ----
    const bool divw = J.op == ARWEN_SHT_OP_DIVERGENCE_FROM_WAIST;
    const int ncol = vec ? 4 : 2;
    js->ncols = ncol * n;
    if (prepare) js->waist_a.alloc((size_t)js->ncols * P->ldw * 4);
    if (prepare && divw) js->waist_b.alloc((size_t)js->ncols * P->ldw * 4);
    if (prepare) js->packed_out.alloc((size_t)js->ncols * P->ldp * 4);
    if (from_waist(J.op)) {
      // a reference waist (*lead, nlat, T+1), already divided by nlon
      sht::WaistFoldArgs wf{};
      wf.w0 = static_cast<const float2*>(J.in0);
      wf.w1 = static_cast<const float2*>(J.in1);
      wf.waist = js->waist_a.as<float>();
      wf.waist1 = divw ? js->waist_b.as<float>() : nullptr;
      wf.rowscale = (vec || divw) ? P->scale_waist_vec.as<float>() : P->scale_waist.as<float>();
      wf.T = T;
      wf.nlat = P->nlat;
      wf.JH = P->JH;
      wf.JHP = P->JHP;
      wf.ldw = P->ldw;
      wf.nfields = n;
      wf.col0 = 0;
      wf.vector = divw ? 2 : vec ? 1 : 0;
      B->waist_folds.push_back(wf);
    }
    sht::FoldArgs fa{};
    fa.fft = P->fft;
    fa.T = T;
    fa.nlat = P->nlat;
    fa.JH = P->JH;
    fa.JHP = P->JHP;
    fa.rp = vec ? P->rp_fold_vec : P->rp_fold;
    fa.ngroups = P->JHP / fa.rp;
    fa.gpb = P->fft_gpb;
    fa.nchunks = (fa.ngroups + fa.gpb - 1) / fa.gpb;
    fa.job.grid0 = static_cast<const float*>(J.in0);
    fa.job.grid1 = static_cast<const float*>(J.in1);
    fa.job.waist0 = js->waist_a.as<float>();
    fa.job.rowscale = vec ? P->scale_fold_vec.as<float>() : P->scale_fold.as<float>();
    fa.job.nfields = n;
    fa.job.col0 = 0;
    fa.job.ldw = P->ldw;
    fa.job.vector = vec ? 1 : 0;
    fa.job.goff = J.grid_offset;
    const bool low = J.op == ARWEN_SHT_OP_VORDIV_FROM_UV && P->m0 > 0;
    // without FP64 low orders (plain FP32) on a wave-resident grid: the
    // recurrence form from m = 4 on, its weights 1/cos(phi) in the fold
    const bool recp = J.op == ARWEN_SHT_OP_VORDIV_FROM_UV && P->m0 == 0 && P->fft.wave_L > 0;
    if (recp) {
      fa.job.rowscale_hi = P->scale_fold_vec_cos.as<float>();
      fa.job.m_hi = sht::kMT;
    }
    if (low) {
      // FP64 folds of u and v, one field each: i U to level columns 2, 3 and
      // i V to 0, 1 (the [iV, iU] arrangement), orders m < m0 also in FP64
      if (prepare) js->low.alloc((size_t)n * 2 * P->m0 * 2 * P->JHP * sizeof(double2));
      for (int f = 0; f < 2; ++f) {
        sht::FoldArgs fh = fa;
        fh.job.grid0 = static_cast<const float*>(f == 0 ? J.in0 : J.in1);
        fh.job.grid1 = nullptr;
        fh.job.vector = 0;
        fh.job.col0 = f == 0 ? 2 : 0;
        fh.job.fstride = 4;
        fh.job.low = js->low.as<double2>() + (size_t)f * P->m0 * 2 * P->JHP;
        fh.job.low_stride = 2 * P->m0 * 2 * P->JHP;
        // both parts in the recurrence form: weights 2 pi w / (a nlon cos(phi))
        fh.job.rowscale64 = P->scale_fold_vec64_cos.as<double>();
        fh.job.rowscale64_hi = P->scale_fold_vec64_cos.as<double>();
        fh.job.m0 = P->m0;
        B->folds.push_back(fh);
      }
      if (prepare) B->fft_bytes += 2.0 * n * fftb;
      if (prepare) js->low_ab.alloc((size_t)n * P->m0 * 2 * (T + 2) * sizeof(double2));
      sht::LowAnaArgs la{};
      la.low = js->low.as<double2>();
      la.p = P->low_ana[kP].as<double>();
      la.ab = js->low_ab.as<double2>();
      la.vort = static_cast<float2*>(J.out0);
      la.div = static_cast<float2*>(J.out1);
      la.T = T;
      la.JH = P->JH;
      la.JHP = P->JHP;
      la.m0 = P->m0;
      la.nlev = n;
      B->low_anas.push_back(la);
    } else if (!from_waist(J.op)) {
      B->folds.push_back(fa);
      if (prepare) B->fft_bytes += (vec ? 2.0 : 1.0) * n * fftb;
    }
    const DataRef wst{js->waist_a.as<char>()};
    for (int mb = low ? P->m0 / sht::kMT : 0; prepare && mb < P->MB; ++mb)
      for (int q = 0; q < 2; ++q) {
        if (low || (recp && mb >= 1)) {
          // recurrence form: [iB, iA] = P^T [iV, iU] / cos(phi); unpack_vector_rec_kernel
          AnaSeg sg[1] = {{kP, 0, wst, 0}};
          add_ana_group(ana, P, mb, q, js->ncols, js->packed_out.as<float>(), sg, 1);
        } else if (vec) {
          // zeta, D = Pm^T [iV, iU]_q + dP^T [U, -V]_{1-q}
          AnaSeg sg[2] = {{kPm, 0, wst, 0}, {kD, 1, wst, kMapAnaVec}};
          add_ana_group(ana, P, mb, q, js->ncols, js->packed_out.as<float>(), sg, 2);
        } else if (divw) {
          // D = Pm^T (iU)_q + dP^T (-V)_{1-q}
          AnaSeg sg[2] = {{kPm, 0, wst, 0}, {kD, 1, DataRef{js->waist_b.as<char>()}, 0}};
          add_ana_group(ana, P, mb, q, js->ncols, js->packed_out.as<float>(), sg, 2);
        } else {
          AnaSeg sg[1] = {{kP, 0, wst, 0}};
          add_ana_group(ana, P, mb, q, js->ncols, js->packed_out.as<float>(), sg, 1);
        }
      }
    if (prepare && low) {  // two per level: the GEMM above m0, the FP64 kernel below
      B->legendre_flops += 2.0 * n * pcf;
      B->low_flops += 2.0 * n * pcf_low;
    } else if (prepare && recp) {  // four per level in the first order tile, two above
      B->legendre_flops += 2.0 * n * pcf + 2.0 * n * pcf_t0;
    } else if (prepare) {
      B->legendre_flops += (vec ? 4.0 : divw ? 2.0 : 1.0) * n * pcf;
    }
    sht::UnpackArgs ua{};
    ua.koff = P->d_koff.as<int>();
    ua.T = T;
    ua.tiles = (T + 1 + 31) / 32;
    ua.job.in = js->packed_out.as<float>();
    ua.job.out0 = static_cast<float2*>(J.out0);
    ua.job.out1 = static_cast<float2*>(J.out1);
    ua.job.nfields = n;
    ua.job.col0 = 0;
    ua.job.ldp = P->ldp;
    ua.job.vector = vec ? 1 : 0;
    ua.job.m0 = low ? P->m0 : 0;
    ua.job.rec = low || recp ? 1 : 0;
    ua.job.mdp = recp ? sht::kMT : 0;
    ua.job.off00 = vec ? nullptr : static_cast<const float2*>(J.coeff_offset);
    B->unpacks.push_back(ua);
  }
}

// The launch arguments of every job, rebuilt from the jobs' current buffers.
void bind_jobs(arwen_sht_batch* B) {
  B->packs.clear();
  B->unfolds.clear();
  B->folds.clear();
  B->unpacks.clear();
  B->waist_unfolds.clear();
  B->waist_folds.clear();
  B->low_syns.clear();
  B->low_anas.clear();
  for (auto& js : B->jobs) setup_job(B, js.get(), false);
}

void plan_batch(arwen_sht_batch* B, const arwen_sht_job* jobs, int njobs) {
  if (njobs <= 0 || !jobs) fail(ARWEN_SHT_EINVAL, "a batch needs at least one job");
  hip_check(hipGetDevice(&B->device), "hipGetDevice");
  B->precision = jobs[0].plan ? jobs[0].plan->precision : 0;
  for (int j = 0; j < njobs; ++j) {
    const arwen_sht_job& J = jobs[j];
    const arwen_sht_plan* P = J.plan;
    if (!P) fail(ARWEN_SHT_EINVAL, "job " + std::to_string(j) + " has no plan");
    if (P->precision != B->precision)
      fail(ARWEN_SHT_EINVAL, "all plans of a batch must share one precision");
    if (P->device != B->device)
      fail(ARWEN_SHT_EINVAL, "job " + std::to_string(j) + "'s plan belongs to another device");
    if (J.count < 0) fail(ARWEN_SHT_EINVAL, "job count must be >= 0");
    const bool vec = J.op == ARWEN_SHT_OP_UV_FROM_VORDIV || J.op == ARWEN_SHT_OP_VORDIV_FROM_UV ||
                     J.op == ARWEN_SHT_OP_VORDIV_FROM_WAIST;
    const bool sgrad = J.op == ARWEN_SHT_OP_SYNTHESIS_GRADIENT;
    const bool two_out = vec || J.op == ARWEN_SHT_OP_GRADIENT || sgrad;
    const bool two_in = vec || J.op == ARWEN_SHT_OP_DIVERGENCE_FROM_WAIST;
    if (J.op < ARWEN_SHT_OP_SYNTHESIS || J.op > ARWEN_SHT_OP_DIVERGENCE_FROM_WAIST)
      fail(ARWEN_SHT_EINVAL, "unknown op " + std::to_string(J.op));
    if (J.count > 0 && (!J.in0 || !J.out0 || (two_in && !J.in1) || (two_out && !J.out1) ||
                        (sgrad && !J.out2)))
      fail(ARWEN_SHT_EINVAL, "job " + std::to_string(j) + " is missing a buffer");
    if (J.flags & ~(ARWEN_SHT_JOB_EXP | ARWEN_SHT_JOB_INVLAP))
      fail(ARWEN_SHT_EINVAL, "job " + std::to_string(j) + " has unknown flags");
    if ((J.flags & ARWEN_SHT_JOB_EXP) && J.op != ARWEN_SHT_OP_SYNTHESIS)
      fail(ARWEN_SHT_EINVAL, "ARWEN_SHT_JOB_EXP applies to ARWEN_SHT_OP_SYNTHESIS only");
    if ((J.flags & ARWEN_SHT_JOB_INVLAP) && (J.op != ARWEN_SHT_OP_SYNTHESIS || J.coeff_offset))
      fail(ARWEN_SHT_EINVAL,
           "ARWEN_SHT_JOB_INVLAP applies to ARWEN_SHT_OP_SYNTHESIS without a coeff_offset");
    if (J.grid_offset && J.op != ARWEN_SHT_OP_ANALYSIS)
      fail(ARWEN_SHT_EINVAL, "grid_offset applies to ARWEN_SHT_OP_ANALYSIS only");
    if (J.coeff_offset && J.op != ARWEN_SHT_OP_ANALYSIS && J.op != ARWEN_SHT_OP_SYNTHESIS &&
        !sgrad)
      fail(ARWEN_SHT_EINVAL,
           "coeff_offset applies to ARWEN_SHT_OP_ANALYSIS, _SYNTHESIS and _SYNTHESIS_GRADIENT only");
    auto js = std::make_unique<JobState>();
    js->job = J;
    setup_job(B, js.get(), true);
    B->jobs.push_back(std::move(js));
  }
  for (int d = 0; d < 2; ++d) build_lists(d, B->groups[d], B->config.cols[d], B->lists[d]);
  if (!g_setup_stream) hip_check(hipDeviceSynchronize(), "batch setup");
}

// plan_batch with its zero fills and uploads on `s`, before whatever the
// caller queues there next; nothing waits for them on the host.
void plan_batch_on(arwen_sht_batch* B, const arwen_sht_job* jobs, int njobs, hipStream_t s) {
  struct Scope {
    Scope(hipStream_t st, std::vector<std::vector<char>>* keep) {
      g_setup_stream = st;
      g_setup_keep = keep;
    }
    ~Scope() {
      g_setup_stream = nullptr;
      g_setup_keep = nullptr;
    }
  } scope(s, &B->setup_keep);
  plan_batch(B, jobs, njobs);
}

// Persistent grid of direction d: a multiple of the XCD count (see
// build_lists); 0 in the config = 8 workgroups per CU.
int gemm_grid(const arwen_sht_batch* B, int d) {
  int g = B->config.grid[d] > 0 ? B->config.grid[d] : 256 * 8;
  g = std::min(g, std::max(B->lists[d].ntiles, 1));
  if (g >= kXcds) g -= g % kXcds;
  return std::max(g, 1);
}

hipError_t launch_gemm(const arwen_sht_batch* B, int d, hipStream_t s) {
  const GemmLists& L = B->lists[d];
  sht::GemmLaunch g{};
  g.mode = B->precision;
  g.dir = d;
  g.bn = L.bn;
  g.acc2 = B->config.split_acc;
  g.panels = L.panels.as<sht::GemmPanel>();
  g.steps = L.steps.as<sht::GemmStep>();
  g.tiles = L.tiles.as<int2>();
  g.ntiles = L.ntiles;
  g.grid = gemm_grid(B, d);
  return sht_launch_legendre_gemm(&g, s);
}

// Phases of a batch run, in order; `ev` (optional) gets one event before
// each phase and one after the last.
// pack, synthesis GEMM, its FP64 low orders, inverse FFT, forward FFT, analysis
// GEMM, its FP64 low orders, unpack
constexpr int kPhases = 8;

void run_batch(arwen_sht_batch* B, hipStream_t s, hipEvent_t* ev = nullptr) {
  int dev;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev != B->device) fail(ARWEN_SHT_EINVAL, "batch used on a device other than its own");
  auto mark = [&](int k) {
    if (ev) hip_check(hipEventRecord(ev[k], s), "hipEventRecord");
  };
  mark(0);
  for (auto& a : B->packs) hip_check(sht_launch_pack(&a, s), "pack launch");
  mark(1);
  hip_check(launch_gemm(B, 0, s), "synthesis GEMM launch");
  mark(2);
  for (auto& a : B->low_syns) hip_check(sht_launch_low_synthesis(&a, s), "low-order synthesis launch");
  mark(3);
  for (auto& a : B->unfolds) hip_check(sht_launch_irfft_unfold(&a, s), "irfft launch");
  for (auto& a : B->waist_unfolds) hip_check(sht_launch_waist_unfold(&a, s), "waist unfold launch");
  mark(4);
  for (auto& a : B->folds) hip_check(sht_launch_rfft_fold(&a, s), "rfft launch");
  for (auto& a : B->waist_folds) hip_check(sht_launch_waist_fold(&a, s), "waist fold launch");
  mark(5);
  hip_check(launch_gemm(B, 1, s), "analysis GEMM launch");
  mark(6);
  for (auto& a : B->low_anas) hip_check(sht_launch_low_analysis(&a, s), "low-order analysis launch");
  mark(7);
  for (auto& a : B->unpacks) hip_check(sht_launch_unpack(&a, s), "unpack launch");
  mark(8);
}

void drop_graph(arwen_sht_batch* B) {
  if (B->exec) {
    (void)hipGraphExecDestroy(B->exec);
    B->exec = nullptr;
  }
  if (B->graph) {
    (void)hipGraphDestroy(B->graph);
    B->graph = nullptr;
  }
}

void apply_config(arwen_sht_batch* B, const arwen_sht_gemm_config& c) {
  for (int d = 0; d < 2; ++d) {
    if (c.cols[d] != 32 && c.cols[d] != 64 && c.cols[d] != 96)
      fail(ARWEN_SHT_EINVAL, "cols must be 32, 64 or 96");
    // the 96-column tile has no split accumulators; in the split modes it
    // would change results, so it is offered only where they are off
    if (c.cols[d] == 96 && c.split_acc && B->precision != ARWEN_SHT_FP32)
      fail(ARWEN_SHT_EINVAL, "cols 96 needs split_acc = 0 in the split-BF16 modes");
    if (c.grid[d] < 0) fail(ARWEN_SHT_EINVAL, "grid must be >= 0");
  }
  if (c.split_acc != 0 && c.split_acc != 1) fail(ARWEN_SHT_EINVAL, "split_acc must be 0 or 1");
  // the lists are rebuilt while no launch of this batch is in flight
  hip_check(hipDeviceSynchronize(), "config change");
  drop_graph(B);
  for (int d = 0; d < 2; ++d)
    if (B->lists[d].bn != c.cols[d]) build_lists(d, B->groups[d], c.cols[d], B->lists[d]);
  B->config = c;
  hip_check(hipDeviceSynchronize(), "config upload");
}

// A single operation runs as a one-job batch kept by the plan per (op,
// count, stream): its workspaces and GEMM work lists depend on those only, so
// a later call with other buffers rebinds the launch arguments on the host
// and launches, with no allocation and no synchronization.  The first call of
// a key allocates, then zero-fills and uploads on the caller's stream ahead
// of its kernels (no host wait); at most kSingleCache keys are kept, and
// evicting one waits for its stream.  Calls of one plan on one stream share
// that key's workspaces in stream order.
constexpr size_t kSingleCache = 16;

int single_job(arwen_sht_plan* P, const arwen_sht_job& job, void* stream) {
  return guarded([&] {
    if (!P) fail(ARWEN_SHT_EINVAL, "plan is NULL");
    const int op = job.op, count = job.count;
    if (count < 0) fail(ARWEN_SHT_EINVAL, "count must be >= 0");
    if (count == 0) return;
    std::lock_guard<std::mutex> lock(P->cache_mu);
    const auto key = std::make_tuple(op, count, stream);
    arwen_sht_batch* B = nullptr;
    for (auto it = P->cache.begin(); it != P->cache.end(); ++it) {
      if (it->first == key) {
        B = it->second;
        P->cache.splice(P->cache.begin(), P->cache, it);
        break;
      }
    }
    arwen_sht_job J = job;
    J.plan = P;
    if (!B) {
      auto nb = std::make_unique<arwen_sht_batch>();
      plan_batch_on(nb.get(), &J, 1, static_cast<hipStream_t>(stream));
      B = nb.release();
      P->cache.push_front({key, B});
      while (P->cache.size() > kSingleCache) {
        auto& victim = P->cache.back();
        hip_check(hipStreamSynchronize(static_cast<hipStream_t>(std::get<2>(victim.first))),
                  "single-operation cache eviction");
        delete victim.second;
        P->cache.pop_back();
      }
    } else {
      B->jobs[0]->job = J;
      bind_jobs(B);
    }
    run_batch(B, static_cast<hipStream_t>(stream));
  });
}

int single_op(arwen_sht_plan* P, int op, const void* in0, const void* in1, void* out0,
              void* out1, int count, void* stream) {
  arwen_sht_job J{};
  J.op = op;
  J.count = count;
  J.in0 = in0;
  J.in1 = in1;
  J.out0 = out0;
  J.out1 = out1;
  return single_job(P, J, stream);
}

}  // namespace

// ---------------------------------------------------------------------------
// Exported C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* arwen_sht_last_error(void) { return g_error.c_str(); }

int arwen_sht_plan_create(const arwen_sht_plan_desc* desc, arwen_sht_plan** out) {
  return guarded([&] {
    if (!out) fail(ARWEN_SHT_EINVAL, "out is NULL");
    *out = nullptr;
    auto P = std::make_unique<arwen_sht_plan>();
    build_plan(P.get(), desc);
    *out = P.release();
  });
}

int arwen_sht_plan_destroy(arwen_sht_plan* plan) {
  return guarded([&] {
    if (!plan) return;
    (void)hipDeviceSynchronize();
    delete plan;
  });
}

int arwen_sht_plan_grid(const arwen_sht_plan* P, double* sin_lat, double* weights) {
  return guarded([&] {
    if (!P) fail(ARWEN_SHT_EINVAL, "plan is NULL");
    if (sin_lat) std::memcpy(sin_lat, P->mu.data(), P->nlat * sizeof(double));
    if (weights) std::memcpy(weights, P->w.data(), P->nlat * sizeof(double));
  });
}

size_t arwen_sht_plan_table_bytes(const arwen_sht_plan* P) { return P ? P->table_bytes() : 0; }

int arwen_sht_plan_fp64_orders(const arwen_sht_plan* P) { return P ? P->m0 : 0; }

int arwen_sht_gauss_nodes(int nlat, double* sin_lat, double* weights) {
  return guarded([&] {
    if (nlat < 1) fail(ARWEN_SHT_EINVAL, "nlat must be >= 1, got " + std::to_string(nlat));
    if (!sin_lat || !weights) fail(ARWEN_SHT_EINVAL, "sin_lat and weights are required");
    DevBuf mu((size_t)nlat * 8, false), w((size_t)nlat * 8, false);
    hip_check(sht_launch_gauss_newton(nlat, mu.as<double>(), w.as<double>(), 0),
              "gauss_newton launch");
    hip_check(hipMemcpy(sin_lat, mu.p, (size_t)nlat * 8, hipMemcpyDeviceToHost), "read nodes");
    hip_check(hipMemcpy(weights, w.p, (size_t)nlat * 8, hipMemcpyDeviceToHost), "read weights");
  });
}

int arwen_sht_plan_table_f64(const arwen_sht_plan* P, int kind, double* out) {
  return guarded([&] {
    if (!P || !out) fail(ARWEN_SHT_EINVAL, "plan and out are required");
    if (kind < 0 || kind > 2) fail(ARWEN_SHT_EINVAL, "kind must be 0 (P), 1 (dP) or 2 (P*m/cos)");
    DevBuf tab((size_t)P->tab_elems * 8, true);
    sht::TableOut t{};
    t.elems = P->tab_elems;
    t.f64_syn[kind] = tab.as<double>();
    hip_check(sht_launch_legendre_tables(P->T, P->nlat, P->JH, P->JHP, P->d_mu.as<double>(),
                                         P->d_koff.as<int>(), &t, 0),
              "legendre_tables launch");
    std::vector<double> h(P->tab_elems);
    hip_check(hipMemcpy(h.data(), tab.p, h.size() * 8, hipMemcpyDeviceToHost), "read table");
    const int T1 = P->T + 1;
    std::fill(out, out + (size_t)T1 * T1 * P->JH, 0.0);
    for (int m = 0; m <= P->T; ++m)
      for (int n = m; n <= P->T; ++n) {
        const int p = (n - m) & 1, idx = (n - m) >> 1;
        const size_t base = (size_t)P->koff[2 * m + p] * P->JHP;
        for (int i = 0; i < P->JH; ++i)
          out[((size_t)m * T1 + n) * P->JH + i] = h[base + (size_t)i * P->kp(m, p) + idx];
      }
  });
}

int arwen_sht_synthesis_to_waist(arwen_sht_plan* P, const void* coeff, void* waist, int count,
                                 int kind, void* s) {
  if (kind < 0 || kind > 2)
    return guarded([&] {
      fail(ARWEN_SHT_EINVAL, "waist kind must be 0 (basis), 1 (basis after i*m) or 2 (dP/dphi), got " +
                                 std::to_string(kind));
    });
  const int op = kind == 0   ? ARWEN_SHT_OP_SYNTHESIS_TO_WAIST
                 : kind == 1 ? ARWEN_SHT_OP_ZONAL_TO_WAIST
                             : ARWEN_SHT_OP_MERIDIONAL_TO_WAIST;
  return single_op(P, op, coeff, nullptr, waist, nullptr, count, s);
}

int arwen_sht_analysis_from_waist(arwen_sht_plan* P, const void* waist, void* coeff, int count,
                                  void* s) {
  return single_op(P, ARWEN_SHT_OP_ANALYSIS_FROM_WAIST, waist, nullptr, coeff, nullptr, count, s);
}

int arwen_sht_divergence_from_waist(arwen_sht_plan* P, const void* uw, const void* vw, void* div,
                                    int count, void* s) {
  return single_op(P, ARWEN_SHT_OP_DIVERGENCE_FROM_WAIST, uw, vw, div, nullptr, count, s);
}

int arwen_sht_vordiv_from_waist(arwen_sht_plan* P, const void* uw, const void* vw, void* vort,
                                void* div, int count, void* s) {
  return single_op(P, ARWEN_SHT_OP_VORDIV_FROM_WAIST, uw, vw, vort, div, count, s);
}

namespace {
int waist_rows(arwen_sht_plan* P, const float* rows_in, float* rows_out, const void* waist_in,
               void* waist_out, int nlead, int nrows, int row0, void* stream, bool to_rows) {
  return guarded([&] {
    if (!P) fail(ARWEN_SHT_EINVAL, "plan is NULL");
    if (nlead < 0 || nrows < 0) fail(ARWEN_SHT_EINVAL, "nlead and nrows must be >= 0");
    if (row0 < 0 || row0 + nrows > P->nlat)
      fail(ARWEN_SHT_EINVAL, "rows [" + std::to_string(row0) + ", " + std::to_string(row0 + nrows) +
                                 ") are outside the waist's " + std::to_string(P->nlat) +
                                 " latitudes");
    if ((long long)nlead * nrows == 0) return;
    if ((long long)nlead * nrows > (long long)INT32_MAX / 64)
      fail(ARWEN_SHT_EINVAL, "nlead x nrows is too large for one call");
    if ((to_rows && (!waist_in || !rows_out)) || (!to_rows && (!rows_in || !waist_out)))
      fail(ARWEN_SHT_EINVAL, "a buffer is NULL");
    int dev;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev != P->device) fail(ARWEN_SHT_EINVAL, "plan used on a device other than its own");
    sht::WaistRowsArgs a{};
    a.fft = P->fft;
    a.T = P->T;
    a.nlat = P->nlat;
    a.rows_in = rows_in;
    a.rows_out = rows_out;
    a.waist_in = static_cast<const float2*>(waist_in);
    a.waist_out = static_cast<float2*>(waist_out);
    a.nlead = nlead;
    a.nrows = nrows;
    a.row0 = row0;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hip_check(to_rows ? sht_launch_rows_from_waist(&a, st) : sht_launch_waist_from_rows(&a, st),
              to_rows ? "rows_from_waist launch" : "waist_from_rows launch");
  });
}
}  // namespace

int arwen_sht_waist_from_rows(arwen_sht_plan* P, const float* rows, void* waist, int nlead,
                              int nrows, int row0, void* s) {
  return waist_rows(P, rows, nullptr, nullptr, waist, nlead, nrows, row0, s, false);
}

int arwen_sht_rows_from_waist(arwen_sht_plan* P, const void* waist, float* rows, int nlead,
                              int nrows, int row0, void* s) {
  return waist_rows(P, nullptr, rows, waist, nullptr, nlead, nrows, row0, s, true);
}

int arwen_sht_analysis(arwen_sht_plan* P, const float* grid, void* coeff, int count, void* s) {
  return single_op(P, ARWEN_SHT_OP_ANALYSIS, grid, nullptr, coeff, nullptr, count, s);
}
int arwen_sht_synthesis(arwen_sht_plan* P, const void* coeff, float* grid, int count, void* s) {
  return single_op(P, ARWEN_SHT_OP_SYNTHESIS, coeff, nullptr, grid, nullptr, count, s);
}
int arwen_sht_zonal_derivative(arwen_sht_plan* P, const void* coeff, float* grid, int count,
                               void* s) {
  return single_op(P, ARWEN_SHT_OP_ZONAL_DERIVATIVE, coeff, nullptr, grid, nullptr, count, s);
}
int arwen_sht_meridional_derivative(arwen_sht_plan* P, const void* coeff, float* grid,
                                    int count, void* s) {
  return single_op(P, ARWEN_SHT_OP_MERIDIONAL_DERIVATIVE, coeff, nullptr, grid, nullptr, count,
                   s);
}
int arwen_sht_gradient(arwen_sht_plan* P, const void* coeff, float* east, float* north,
                       int count, void* s) {
  return single_op(P, ARWEN_SHT_OP_GRADIENT, coeff, nullptr, east, north, count, s);
}
int arwen_sht_synthesis_gradient(arwen_sht_plan* P, const void* coeff, float* grid, float* east,
                                 float* north, int count, void* s) {
  arwen_sht_job J{};
  J.op = ARWEN_SHT_OP_SYNTHESIS_GRADIENT;
  J.count = count;
  J.in0 = coeff;
  J.out0 = grid;
  J.out1 = east;
  J.out2 = north;
  return single_job(P, J, s);
}

int arwen_sht_analysis_about(arwen_sht_plan* P, const float* grid, void* coeff, int count,
                             const float* grid_offset, const void* coeff_offset, void* s) {
  arwen_sht_job J{};
  J.op = ARWEN_SHT_OP_ANALYSIS;
  J.count = count;
  J.in0 = grid;
  J.out0 = coeff;
  J.grid_offset = grid_offset;
  J.coeff_offset = coeff_offset;
  return single_job(P, J, s);
}

int arwen_sht_synthesis_about(arwen_sht_plan* P, const void* coeff, float* grid, int count,
                              const void* coeff_offset, int flags, void* s) {
  arwen_sht_job J{};
  J.op = ARWEN_SHT_OP_SYNTHESIS;
  J.count = count;
  J.in0 = coeff;
  J.out0 = grid;
  J.coeff_offset = coeff_offset;
  J.flags = flags;
  return single_job(P, J, s);
}

int arwen_sht_uv_from_vordiv(arwen_sht_plan* P, const void* vort, const void* div, float* u,
                             float* v, int count, void* s) {
  return single_op(P, ARWEN_SHT_OP_UV_FROM_VORDIV, vort, div, u, v, count, s);
}
int arwen_sht_vordiv_from_uv(arwen_sht_plan* P, const float* u, const float* v, void* vort,
                             void* div, int count, void* s) {
  return single_op(P, ARWEN_SHT_OP_VORDIV_FROM_UV, u, v, vort, div, count, s);
}

int arwen_sht_spectral_op(arwen_sht_plan* P, int op, const void* in, void* out, int count,
                          void* s) {
  return guarded([&] {
    if (!P) fail(ARWEN_SHT_EINVAL, "plan is NULL");
    if (op < 0 || op > 2) fail(ARWEN_SHT_EINVAL, "op must be 0 (project), 1 or 2");
    if (count < 0) fail(ARWEN_SHT_EINVAL, "count must be >= 0");
    if (count == 0) return;
    if (!in || !out) fail(ARWEN_SHT_EINVAL, "in and out are required");
    long long total = (long long)count * (P->T + 1) * (P->T + 1);
    hip_check(sht_launch_spectral_elementwise(static_cast<const float2*>(in),
                                              static_cast<float2*>(out), total, P->T, op,
                                              (float)P->radius, static_cast<hipStream_t>(s)),
              "spectral op launch");
  });
}

namespace {
constexpr int kSampleColPad = 8;  // sampler field columns are padded to the widest field tile

int sample_columns(int count, int kind) {
  const long long cols = (kind == 2 ? 2LL : 1LL) * count;
  return (int)((cols + kSampleColPad - 1) / kSampleColPad * kSampleColPad);
}

// The point sampler's recurrence coefficients per packed (n, m) and per order,
// with the reference's numpy expressions (sampling.py:83-112) in double.
void build_sample_tables(arwen_sht_plan* P) {
  const int T = P->T;
  const long long K = (long long)(T + 1) * (T + 2) / 2;
  std::vector<double> rec(3 * K, 0.0), diag(2 * (T + 1), 0.0);
  long long k = 0;
  for (int m = 0; m <= T; ++m)
    for (int n = m; n <= T; ++n, ++k) {
      const double nn = n, mm2 = (double)((long long)m * m), nm2 = (double)((long long)n * n - (long long)m * m);
      if (n >= m + 2) {
        rec[k] = std::sqrt((4.0 * nn * nn - 1.0) / nm2);
        rec[K + k] = std::sqrt(((nn - 1.0) * (nn - 1.0) - mm2) / (4.0 * (nn - 1.0) * (nn - 1.0) - 1.0));
      }
      if (n > m) rec[2 * K + k] = std::sqrt(nm2 * (2.0 * nn + 1.0) / (2.0 * nn - 1.0));
    }
  for (int m = 1; m <= T; ++m) diag[m] = -std::sqrt((2 * m + 1) / (2.0 * m));
  for (int m = 0; m <= T; ++m) diag[T + 1 + m] = std::sqrt((double)(2 * m + 3));
  // both uploads succeed before the plan holds either (the sampler checks both)
  DevBuf r, dg;
  upload(r, rec);
  upload(dg, diag);
  std::swap(P->sample_rec.p, r.p);
  std::swap(P->sample_rec.bytes, r.bytes);
  std::swap(P->sample_diag.p, dg.p);
  std::swap(P->sample_diag.bytes, dg.bytes);
}
}  // namespace

// Order-range parts per (point block, field tile): a function of T only, so
// a point's arithmetic never depends on how many points share a launch.
sht::SampleCuts sample_cuts(int T) {
  sht::SampleCuts c{};
  c.parts = std::max(1, std::min(sht::kSampleMaxParts, (T + 1) / 16));
  const long long K = (long long)(T + 1) * (T + 2) / 2;
  c.m[0] = 0;
  long long acc = 0;
  int r = 1;
  for (int m = 0; m <= T && r < c.parts; ++m) {
    acc += T + 1 - m;  // entries of orders 0..m
    if (acc * c.parts >= (long long)r * K) c.m[r++] = m + 1;
  }
  for (; r < c.parts; ++r) c.m[r] = c.m[r - 1] + 1;  // not reached for T >= 15
  c.m[c.parts] = T + 1;
  return c;
}

size_t arwen_sht_sample_workspace_bytes(const arwen_sht_plan* P, int count, int kind,
                                        int npoints) {
  if (!P || count < 0 || kind < 0 || kind > 2 || npoints < 0) return 0;
  const long long K = (long long)(P->T + 1) * (P->T + 2) / 2;
  const int ldc = sample_columns(count, kind);
  const int R = sample_cuts(P->T).parts;
  size_t bytes = (size_t)K * ldc * sizeof(double2);  // packed coefficients, either precision
  // per-part raw sums: one plane per part for values, two for derivatives
  if (R > 1) bytes += (size_t)(kind == 0 ? 1 : 2) * R * ldc * (size_t)npoints * sizeof(double);
  return bytes;
}

namespace {
// The sampler's argument checks; returns false when there is nothing to do.
bool sample_checks(const arwen_sht_plan* P, const void* in0, const void* in1, int count,
                   const double* sin_lat, const double* cos_lat, const double* lon_rad,
                   int npoints, int kind, int precision, const double* out0, const double* out1) {
  if (!P) fail(ARWEN_SHT_EINVAL, "plan is NULL");
  if (kind < 0 || kind > 2)
    fail(ARWEN_SHT_EINVAL, "sample kind must be 0 (values), 1 (gradient) or 2 (wind), got " +
                               std::to_string(kind));
  if (precision != 0 && precision != 1)
    fail(ARWEN_SHT_EINVAL, "sample precision must be 0 (float64) or 1 (float32), got " +
                               std::to_string(precision));
  if (count < 0 || npoints < 0) fail(ARWEN_SHT_EINVAL, "count and npoints must be >= 0");
  if ((long long)count * npoints == 0) return false;
  if (!in0 || (kind == 2 && !in1)) fail(ARWEN_SHT_EINVAL, "an input field is NULL");
  if (!sin_lat || !lon_rad || (kind != 0 && !cos_lat))
    fail(ARWEN_SHT_EINVAL, "a point coordinate array is NULL");
  if (!out0 || (kind != 0 && !out1)) fail(ARWEN_SHT_EINVAL, "an output is NULL");
  int dev;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev != P->device) fail(ARWEN_SHT_EINVAL, "plan used on a device other than its own");
  if (!P->sample_rec.p || !P->sample_diag.p)
    fail(ARWEN_SHT_EHIP, "the plan's sampler tables are missing");
  return true;
}

// The pack and sampler launches with the workspace at `workspace` (at least
// arwen_sht_sample_workspace_bytes of the call); allocates nothing.
void sample_launch(arwen_sht_plan* P, const void* in0, const void* in1, int count,
                   const double* sin_lat, const double* cos_lat, const double* lon_rad,
                   int npoints, int kind, int precision, void* workspace, double* out0,
                   double* out1, hipStream_t st);
}  // namespace

int arwen_sht_sample_static(arwen_sht_plan* P, const void* in0, const void* in1, int count,
                            const double* sin_lat, const double* cos_lat, const double* lon_rad,
                            int npoints, int kind, int precision, double* out0, double* out1,
                            void* s) {
  return guarded([&] {
    if (!sample_checks(P, in0, in1, count, sin_lat, cos_lat, lon_rad, npoints, kind, precision,
                       out0, out1))
      return;
    const size_t need = arwen_sht_sample_workspace_bytes(P, count, kind, npoints);
    if (need > ARWEN_SHT_SAMPLE_STATIC_BYTES)
      fail(ARWEN_SHT_EINVAL, "the sample needs a " + std::to_string(need) +
                                 "-byte workspace, more than the static array's " +
                                 std::to_string((size_t)ARWEN_SHT_SAMPLE_STATIC_BYTES) +
                                 " bytes: use arwen_sht_sample or fewer points");
    void* scratch = sht_static_scratch(P->device);
    if (!scratch) fail(ARWEN_SHT_EHIP, "the static sampler workspace of this device is unresolved");
    sample_launch(P, in0, in1, count, sin_lat, cos_lat, lon_rad, npoints, kind, precision,
                  scratch, out0, out1, static_cast<hipStream_t>(s));
  });
}

int arwen_sht_sample(arwen_sht_plan* P, const void* in0, const void* in1, int count,
                     const double* sin_lat, const double* cos_lat, const double* lon_rad,
                     int npoints, int kind, int precision, void* workspace,
                     size_t workspace_bytes, double* out0, double* out1, void* s) {
  return guarded([&] {
    if (!sample_checks(P, in0, in1, count, sin_lat, cos_lat, lon_rad, npoints, kind, precision,
                       out0, out1))
      return;
    const size_t need = arwen_sht_sample_workspace_bytes(P, count, kind, npoints);
    if (!workspace || workspace_bytes < need)
      fail(ARWEN_SHT_EINVAL, "sample workspace must hold " + std::to_string(need) + " bytes");
    sample_launch(P, in0, in1, count, sin_lat, cos_lat, lon_rad, npoints, kind, precision,
                  workspace, out0, out1, static_cast<hipStream_t>(s));
  });
}

namespace {
void sample_launch(arwen_sht_plan* P, const void* in0, const void* in1, int count,
                   const double* sin_lat, const double* cos_lat, const double* lon_rad,
                   int npoints, int kind, int precision, void* workspace, double* out0,
                   double* out1, hipStream_t st) {
  {
    const int ldc = sample_columns(count, kind);
    sht::SamplePackArgs pa{};
    pa.in0 = static_cast<const float2*>(in0);
    pa.in1 = static_cast<const float2*>(in1);
    pa.out = workspace;
    pa.f64 = precision == 0 ? 1 : 0;
    pa.radius = (double)(float)P->radius;  // the spectral operator kernel's lambda_n
    pa.T = P->T;
    pa.count = count;
    pa.ldc = ldc;
    pa.wind = kind == 2 ? 1 : 0;
    hip_check(sht_launch_sample_pack(&pa, st), "sample pack launch");
    sht::SampleArgs a{};
    a.coef = workspace;
    a.rec = P->sample_rec.as<double>();
    a.diag = P->sample_diag.as<double>();
    a.sin_lat = sin_lat;
    a.cos_lat = cos_lat ? cos_lat : sin_lat;  // unread for values
    a.lon = lon_rad;
    a.out0 = out0;
    a.out1 = out1 ? out1 : out0;  // unwritten for values
    a.cuts = sample_cuts(P->T);
    const long long K = (long long)(P->T + 1) * (P->T + 2) / 2;
    a.partial = a.cuts.parts > 1
                    ? reinterpret_cast<double*>(static_cast<char*>(workspace) +
                                                (size_t)K * ldc * sizeof(double2))
                    : nullptr;
    a.radius = P->radius;
    a.T = P->T;
    a.npts = npoints;
    a.ldc = ldc;
    a.nout = count;
    a.kind = kind;
    a.f32 = precision;
    hip_check(sht_launch_sample(&a, st), "sample launch");
  }
}
}  // namespace

int arwen_sht_vertical_solve(arwen_sht_plan* P, int nlev, const void* div, const void* theta,
                             const void* lnps, const float* b_matrix, const float* c_matrix,
                             const float* m_matrix, const float* inverse, const float* k2,
                             float s_bx, float s_md, float s_c, float s_cd, float s_cn,
                             void* div_out, void* theta_out, void* lnps_out, void* s) {
  return guarded([&] {
    if (!P) fail(ARWEN_SHT_EINVAL, "plan is NULL");
    if (nlev < 1 || nlev > sht::kMaxLevels)
      fail(ARWEN_SHT_EINVAL, "nlev must be in 1.." + std::to_string(sht::kMaxLevels) + ", got " +
                                 std::to_string(nlev));
    if (!div || !theta || !lnps || !div_out || !theta_out || !lnps_out)
      fail(ARWEN_SHT_EINVAL, "a spectral input or output is NULL");
    if (!b_matrix || !c_matrix || !inverse || !k2)
      fail(ARWEN_SHT_EINVAL, "b_matrix, c_matrix, inverse and k2 are required");
    if (s_md != 0.f && !m_matrix) fail(ARWEN_SHT_EINVAL, "s_md != 0 needs m_matrix");
    int dev;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev != P->device) fail(ARWEN_SHT_EINVAL, "plan used on a device other than its own");
    sht::VerticalArgs a{};
    a.div = static_cast<const float2*>(div);
    a.theta = static_cast<const float2*>(theta);
    a.lnps = static_cast<const float2*>(lnps);
    a.b_matrix = b_matrix;
    a.c_matrix = c_matrix;
    a.m_matrix = s_md != 0.f ? m_matrix : nullptr;
    a.inverse = inverse;
    a.k2 = k2;
    a.d_out = static_cast<float2*>(div_out);
    a.theta_out = static_cast<float2*>(theta_out);
    a.lnps_out = static_cast<float2*>(lnps_out);
    a.s_bx = s_bx;
    a.s_md = s_md;
    a.s_c = s_c;
    a.s_cd = s_cd;
    a.s_cn = s_cn;
    a.T = P->T;
    a.nlev = nlev;
    a.mode = 0;
    hip_check(sht_launch_vertical_solve(&a, static_cast<hipStream_t>(s)), "vertical solve launch");
  });
}

namespace {
void vertical_common(arwen_sht_plan* P, int nlev, const void* div, const void* theta,
                     const void* lnps, const float* b, const float* c, const float* k2,
                     void* div_out, void* theta_out, void* lnps_out, sht::VerticalArgs& a) {
  if (!P) fail(ARWEN_SHT_EINVAL, "plan is NULL");
  if (nlev < 1 || nlev > sht::kMaxLevels)
    fail(ARWEN_SHT_EINVAL, "nlev must be in 1.." + std::to_string(sht::kMaxLevels) + ", got " +
                               std::to_string(nlev));
  if (!div || !theta || !lnps || !div_out || !theta_out || !lnps_out)
    fail(ARWEN_SHT_EINVAL, "a spectral input or output is NULL");
  if (!b || !c || !k2) fail(ARWEN_SHT_EINVAL, "b_matrix, c_matrix and k2 are required");
  int dev;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev != P->device) fail(ARWEN_SHT_EINVAL, "plan used on a device other than its own");
  a.div = static_cast<const float2*>(div);
  a.theta = static_cast<const float2*>(theta);
  a.lnps = static_cast<const float2*>(lnps);
  a.b_matrix = b;
  a.c_matrix = c;
  a.k2 = k2;
  a.d_out = static_cast<float2*>(div_out);
  a.theta_out = static_cast<float2*>(theta_out);
  a.lnps_out = static_cast<float2*>(lnps_out);
  a.T = P->T;
  a.nlev = nlev;
}
}  // namespace

int arwen_sht_vertical_tendencies(arwen_sht_plan* P, int nlev, const void* div, const void* theta,
                                  const void* lnps, const float* b_matrix, const float* c_matrix,
                                  const float* k2, float weight, void* div_out, void* theta_out,
                                  void* lnps_out, void* s) {
  return guarded([&] {
    sht::VerticalArgs a{};
    vertical_common(P, nlev, div, theta, lnps, b_matrix, c_matrix, k2, div_out, theta_out,
                    lnps_out, a);
    a.s_bx = weight;
    a.s_c = weight;
    a.mode = 1;
    hip_check(sht_launch_vertical_solve(&a, static_cast<hipStream_t>(s)),
              "vertical tendencies launch");
  });
}

int arwen_sht_vertical_half_apply(arwen_sht_plan* P, int nlev, const void* div, const void* theta,
                                  const void* lnps, const float* b_matrix, const float* c_matrix,
                                  const float* p_d, const float* q_d, const float* p_x,
                                  const float* q_x, const float* k2, void* div_out,
                                  void* theta_out, void* lnps_out, void* s) {
  return guarded([&] {
    sht::VerticalArgs a{};
    vertical_common(P, nlev, div, theta, lnps, b_matrix, c_matrix, k2, div_out, theta_out,
                    lnps_out, a);
    if (!p_d || !q_d || !p_x || !q_x) fail(ARWEN_SHT_EINVAL, "the four half maps are required");
    a.p_d = p_d;
    a.q_d = q_d;
    a.p_x = p_x;
    a.q_x = q_x;
    a.mode = 2;
    hip_check(sht_launch_vertical_solve(&a, static_cast<hipStream_t>(s)), "half map launch");
  });
}

int arwen_sht_batch_create(const arwen_sht_job* jobs, int njobs, arwen_sht_batch** out) {
  return guarded([&] {
    if (!out) fail(ARWEN_SHT_EINVAL, "out is NULL");
    *out = nullptr;
    auto B = std::make_unique<arwen_sht_batch>();
    plan_batch(B.get(), jobs, njobs);
    *out = B.release();
  });
}

int arwen_sht_batch_destroy(arwen_sht_batch* B) {
  return guarded([&] {
    if (!B) return;
    (void)hipDeviceSynchronize();
    delete B;
  });
}

int arwen_sht_batch_run(arwen_sht_batch* B, void* stream) {
  return guarded([&] {
    if (!B) fail(ARWEN_SHT_EINVAL, "batch is NULL");
    run_batch(B, static_cast<hipStream_t>(stream));
  });
}

int arwen_sht_batch_capture(arwen_sht_batch* B, void* stream) {
  return guarded([&] {
    if (!B) fail(ARWEN_SHT_EINVAL, "batch is NULL");
    if (B->exec) return;
    hipStream_t s = static_cast<hipStream_t>(stream);
    bool own = false;
    if (s == nullptr) {  // the null stream cannot be captured
      hip_check(hipStreamCreate(&s), "hipStreamCreate");
      own = true;
    }
    hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
    try {
      run_batch(B, s);
    } catch (...) {
      hipGraph_t g;
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      if (own) (void)hipStreamDestroy(s);
      throw;
    }
    hip_check(hipStreamEndCapture(s, &B->graph), "end capture");
    hip_check(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0), "graph instantiate");
    if (own) (void)hipStreamDestroy(s);
  });
}

int arwen_sht_batch_launch(arwen_sht_batch* B, void* stream) {
  int rc = arwen_sht_batch_capture(B, stream);
  if (rc != ARWEN_SHT_OK) return rc;
  return guarded([&] {
    hip_check(hipGraphLaunch(B->exec, static_cast<hipStream_t>(stream)), "graph launch");
  });
}

int arwen_sht_batch_profile(arwen_sht_batch* B, void* stream, int reps, double* ms) {
  return guarded([&] {
    if (!B || !ms) fail(ARWEN_SHT_EINVAL, "batch and ms are required");
    if (reps < 1) fail(ARWEN_SHT_EINVAL, "reps must be >= 1");
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::vector<hipEvent_t> ev((size_t)(kPhases + 1) * reps, nullptr);
    for (auto& e : ev) hip_check(hipEventCreate(&e), "hipEventCreate");
    try {
      for (int r = 0; r < reps; ++r) run_batch(B, s, &ev[(size_t)r * (kPhases + 1)]);
      hip_check(hipEventSynchronize(ev.back()), "hipEventSynchronize");
      for (int k = 0; k < kPhases; ++k) ms[k] = 0.0;
      for (int r = 0; r < reps; ++r)
        for (int k = 0; k < kPhases; ++k) {
          float t = 0.f;
          hip_check(hipEventElapsedTime(&t, ev[(size_t)r * (kPhases + 1) + k],
                                        ev[(size_t)r * (kPhases + 1) + k + 1]),
                    "hipEventElapsedTime");
          ms[k] += t / reps;
        }
    } catch (...) {
      for (auto& e : ev) (void)hipEventDestroy(e);
      throw;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
  });
}


int arwen_sht_batch_get_config(const arwen_sht_batch* B, arwen_sht_gemm_config* out) {
  return guarded([&] {
    if (!B || !out) fail(ARWEN_SHT_EINVAL, "batch and out are required");
    *out = B->config;
  });
}

int arwen_sht_batch_set_config(arwen_sht_batch* B, const arwen_sht_gemm_config* config) {
  return guarded([&] {
    if (!B || !config) fail(ARWEN_SHT_EINVAL, "batch and config are required");
    apply_config(B, *config);
  });
}

int arwen_sht_batch_autotune(arwen_sht_batch* B, void* stream, int reps,
                             arwen_sht_gemm_config* out) {
  return guarded([&] {
    if (!B) fail(ARWEN_SHT_EINVAL, "batch is NULL");
    if (reps < 1) fail(ARWEN_SHT_EINVAL, "reps must be >= 1");
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    arwen_sht_gemm_config best = B->config;
    try {
      // Each direction is timed on its own, over every column tile and grid;
      // the operands are the batch's own work buffers (their contents do not
      // change the timing).  The split-accumulator choice is a precision
      // setting and is kept as configured; the 96-column tile, which has no
      // split accumulators, is a candidate only where they are off.
      const bool wide = B->precision == ARWEN_SHT_FP32 || !B->config.split_acc;
      const std::vector<int> cols = wide ? std::vector<int>{32, 64, 96} : std::vector<int>{32, 64};
      const int grids[4] = {256 * 4, 256 * 8, 256 * 16, 256 * 32};
      for (int d = 0; d < 2; ++d) {
        if (B->lists[d].ntiles == 0) continue;
        float best_ms = 0.f;
        for (int bn : cols) {
          for (int gr : grids) {
            arwen_sht_gemm_config c = best;
            c.cols[d] = bn;
            c.grid[d] = gr;
            apply_config(B, c);
            hip_check(launch_gemm(B, d, s), "autotune warm-up");
            hip_check(hipEventRecord(e0, s), "hipEventRecord");
            for (int r = 0; r < reps; ++r) hip_check(launch_gemm(B, d, s), "autotune launch");
            hip_check(hipEventRecord(e1, s), "hipEventRecord");
            hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
            float ms = 0.f;
            hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
            if (best_ms == 0.f || ms < best_ms) {
              best_ms = ms;
              best.cols[d] = bn;
              best.grid[d] = gr;
            }
          }
        }
      }
      apply_config(B, best);
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (out) *out = best;
  });
}

int arwen_sht_batch_work(const arwen_sht_batch* B, double* flops, double* bytes,
                         double* low_flops) {
  return guarded([&] {
    if (!B) fail(ARWEN_SHT_EINVAL, "batch is NULL");
    if (flops) *flops = B->legendre_flops;
    if (bytes) *bytes = B->fft_bytes;
    if (low_flops) *low_flops = B->low_flops;
  });
}

}  // extern "C"
