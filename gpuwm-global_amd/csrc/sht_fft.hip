// Longitude real FFTs fused with the parity fold (analysis) and unfold
// (synthesis).
//
// The waist is stored m-tiled:  W[c][m/4][p][i][m%4], per data column c
// (2*field + re/im), fp32 in every precision mode (the split modes cut the
// analysis waist into bf16 pieces as the GEMM loads it).
//
// A block holds `rp` wavefronts; wavefront w owns folded row i0 + w (the
// mirror pair north = nlat-1-i, south = i) and runs its own FFT in its own
// LDS slice, so rp FFTs advance together between the block barriers of the
// Stockham passes.  Neighbouring wavefronts write neighbouring runs of each
// m-tile (4 orders per latitude), which meet in L2 as whole sectors.
//
// Two real rows ride one complex FFT of length N = nlon:
//   z = north + i*south,  X_N[m] = (Z[m] + conj Z[N-m]) / 2,
//                         X_S[m] = (Z[m] - conj Z[N-m]) / (2i),
// and the inverse builds Z[m] = A[m] + iB[m], Z[N-m] = conj A[m] + i conj B[m].
// Reference semantics: analysis waist = rfft(f)[..., :T+1] / nlon
// (reference spectral/transform.py:1386-1402); synthesis = irfft of the
// zero-padded spectrum scaled by nlon (transform.py:1654-1702), i.e.
// f = Re G_0 + 2 Re sum_{m>=1} G_m e^{i m lambda}.
//
// Stockham autosort, radix 4/2/3 butterflies and a generic-radix DFT for any
// other prime factor, twiddles from a table of e^{-2 pi i k/N} computed in
// FP64 and rounded once.
#include <algorithm>
#include <map>
#include <mutex>
#include <type_traits>
#include <utility>

#include "sht_kernels.h"

namespace sht {

constexpr int kWave = 64;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// The same operations in FP64 (the vector operators' FP64 FFTs).
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ float2 cmk(float x, float y) { return make_float2(x, y); }
__device__ __forceinline__ double2 cmk(double x, double y) { return make_double2(x, y); }

// Twiddle e^{sign * 2 pi i k / N} from the forward table.
template <class C>
__device__ __forceinline__ C twiddle(const C* tw, int k, int sign) {
  C t = tw[k];
  return sign < 0 ? t : cconj(t);
}

template <int R>
__device__ __forceinline__ void dft_small(float2* v, int sign) {
  if constexpr (R == 2) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    float2 a2 = cadd(v[1], v[3]), a3 = csub(v[1], v[3]);
    // a3 times -i (forward) or +i (inverse)
    float2 a3r = sign < 0 ? make_float2(a3.y, -a3.x) : make_float2(-a3.y, a3.x);
    v[0] = cadd(a0, a2);
    v[2] = csub(a0, a2);
    v[1] = cadd(a1, a3r);
    v[3] = csub(a1, a3r);
  } else if constexpr (R == 3) {
    const float c = -0.5f;
    const float s = (sign < 0 ? -1.0f : 1.0f) * 0.86602540378443864676f;
    float2 t1 = cadd(v[1], v[2]);
    float2 t2 = csub(v[1], v[2]);
    float2 m1 = make_float2(v[0].x + c * t1.x, v[0].y + c * t1.y);
    float2 m2 = make_float2(-s * t2.y, s * t2.x);  // i*s*t2
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m1, m2);
    v[2] = csub(m1, m2);
  }
}

// One Stockham pass of radix R over one wavefront's N points.
template <int R>
__device__ void stockham_pass(const float2* __restrict__ in, float2* __restrict__ out, int N,
                              int Ns, const float2* tw, int sign, int lane) {
  const int nb = N / R;
  const int tstep = N / (Ns * R);
  for (int j = lane; j < nb; j += kWave) {
    const int k = j % Ns;
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      float2 x = in[j + r * nb];
      v[r] = (r == 0 || k == 0) ? x : cmul(x, twiddle(tw, r * k * tstep, sign));
    }
    dft_small<R>(v, sign);
    const int d = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) out[d + r * Ns] = v[r];
  }
}

__device__ void stockham_pass_generic(const float2* __restrict__ in, float2* __restrict__ out,
                                      int N, int Ns, int R, const float2* tw, int sign,
                                      int lane) {
  const int nb = N / R;
  const int rstep = N / R;  // twiddle stride of the R-point DFT
  const int tstep = N / (Ns * R);
  for (int j = lane; j < nb; j += kWave) {
    const int k = j % Ns;
    const int d = (j / Ns) * Ns * R + k;
    for (int q = 0; q < R; ++q) {
      float2 acc = make_float2(0.f, 0.f);
      for (int r = 0; r < R; ++r) {
        float2 x = in[j + r * nb];
        if (r > 0 && k > 0) x = cmul(x, twiddle(tw, r * k * tstep, sign));
        if (r > 0 && q > 0) x = cmul(x, twiddle(tw, ((r * q) % R) * rstep, sign));
        acc = cadd(acc, x);
      }
      out[d + q * Ns] = acc;
    }
  }
}

// FFT of N points in buf0 (one wavefront's slice); every wavefront of the
// block runs the same pass sequence, so block barriers separate the passes.
// Returns the buffer holding the result.
__device__ float2* wave_fft(float2* buf0, float2* buf1, const FftPlan& plan, const float2* tw,
                            int sign, int lane) {
  float2* in = buf0;
  float2* out = buf1;
  int Ns = 1;
  for (int s = 0; s < plan.nstages; ++s) {
    const int R = plan.radix[s];
    if (R == 4) stockham_pass<4>(in, out, plan.N, Ns, tw, sign, lane);
    else if (R == 2) stockham_pass<2>(in, out, plan.N, Ns, tw, sign, lane);
    else if (R == 3) stockham_pass<3>(in, out, plan.N, Ns, tw, sign, lane);
    else stockham_pass_generic(in, out, plan.N, Ns, R, tw, sign, lane);
    __syncthreads();
    Ns *= R;
    float2* t = in;
    in = out;
    out = t;
  }
  return in;
}

__device__ __forceinline__ size_t waist_index(int c, int ldw, int JHP, int m, int p, int i) {
  return (size_t)c * ldw + ((size_t)((m / kMT) * 2 + p) * JHP + i) * kMT + (m % kMT);
}

// ---------------------------------------------------------------------------
// Analysis: grid rows -> folded, weighted waist.
// ---------------------------------------------------------------------------
// A block walks `gpb` consecutive row groups of one field; each wavefront
// fetches its next mirror pair's rows into registers (PF floats per lane and
// row, PF * 64 >= N) while it transforms the current pair.  PF = 0 reads the
// rows straight into LDS instead.  LDS: the N twiddles, then per wavefront
// two N-point buffers plus a third holding the u spectrum while v is
// transformed (vector jobs).
template <int PF>
__global__ void __launch_bounds__(kFftThreads) rfft_fold_kernel(FoldArgs a) {
  extern __shared__ float2 lds[];
  const int N = a.fft.N;
  const int T1 = a.T + 1;
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  const int chunk = blockIdx.x % a.nchunks;
  const int item = blockIdx.x / a.nchunks;  // field (scalar) or level (vector)
  if (item >= a.job.nfields) return;        // block-uniform
  const int nsrc = a.job.vector ? 2 : 1;
  const int nbuf = a.job.vector ? 3 : 2;
  float2* tw = lds;
  float2* base = lds + N + (size_t)w * nbuf * N;
  for (int k = threadIdx.x; k < N; k += blockDim.x) tw[k] = a.fft.tw[k];
  const size_t plane = (size_t)a.nlat * N;
  const float* gsrc[2] = {a.job.grid0 + (size_t)item * plane,
                          a.job.vector ? a.job.grid1 + (size_t)item * plane : nullptr};
  const int g0 = chunk * a.gpb;
  const int g1 = min(g0 + a.gpb, a.ngroups);
  const int Tpad = ((T1 + kMT - 1) / kMT) * kMT;

  float rn[2][PF > 0 ? PF : 1], rs[2][PF > 0 ? PF : 1];
  auto fetch = [&](int g) {
    if constexpr (PF > 0) {
      const int i = g * a.rp + w;
      if (i >= a.JH) return;
#pragma unroll
      for (int src = 0; src < 2; ++src) {
        if (src >= nsrc) break;
        const float* n_ = gsrc[src] + (size_t)(a.nlat - 1 - i) * N;
        const float* s_ = gsrc[src] + (size_t)i * N;
#pragma unroll
        for (int j = 0; j < PF; ++j) {
          const int k = lane + kWave * j;
          rn[src][j] = k < N ? n_[k] : 0.f;
          rs[src][j] = k < N ? s_[k] : 0.f;
        }
      }
    }
  };
  // a scalar field's grid offset is subtracted as its rows enter the FFT
  // (x - 0 = x exactly, so no offset changes no bit)
  const float goff = (!a.job.vector && a.job.goff) ? a.job.goff[item] : 0.f;
  auto put = [&](int g, int src, float2* b) {
    const int i = g * a.rp + w;
    if (i >= a.JH) return;
    if constexpr (PF > 0) {
#pragma unroll
      for (int j = 0; j < PF; ++j) {
        const int k = lane + kWave * j;
        if (k < N) b[k] = make_float2(rn[src][j] - goff, rs[src][j] - goff);
      }
    } else {
      const float* n_ = gsrc[src] + (size_t)(a.nlat - 1 - i) * N;
      const float* s_ = gsrc[src] + (size_t)i * N;
      for (int k = lane; k < N; k += kWave) b[k] = make_float2(n_[k] - goff, s_[k] - goff);
    }
  };

  fetch(g0);
  for (int g = g0; g < g1; ++g) {
    const int i = g * a.rp + w;
    const bool live = i < a.JH;
    const float s = live ? a.job.rowscale[i] : 0.f;
    put(g, 0, base);
    float vreg[PF > 0 ? PF : 1], vsreg[PF > 0 ? PF : 1];  // v rows wait in registers
    if constexpr (PF > 0) {
#pragma unroll
      for (int j = 0; j < PF; ++j) {
        vreg[j] = rn[1][j];
        vsreg[j] = rs[1][j];
      }
    }
    if (g + 1 < g1) fetch(g + 1);  // next pair's rows load during these FFTs
    __syncthreads();
    float2* Zsrc[2] = {nullptr, nullptr};
    float2* Z = wave_fft(base, base + N, a.fft, tw, -1, lane);
    Zsrc[0] = Z;
    if (nsrc == 2) {
      if (Z != base) {
        // keep the u spectrum in buffer 0 while v is transformed in 1 and 2
        for (int k = lane; k < N; k += kWave) base[k] = Z[k];
        Zsrc[0] = base;
      }
      if (live) {
        if constexpr (PF > 0) {
#pragma unroll
          for (int j = 0; j < PF; ++j) {
            const int k = lane + kWave * j;
            if (k < N) base[N + k] = make_float2(vreg[j], vsreg[j]);
          }
        } else {
          put(g, 1, base + N);
        }
      }
      __syncthreads();
      Zsrc[1] = wave_fft(base + N, base + 2 * N, a.fft, tw, -1, lane);
    }
    // Fold and write: E = (X_N + X_S) s, O = (X_N - X_S) s for m = 0..T,
    // zeros for the padding rows and the padding orders of the last m-tile.
    for (int m = lane; m < Tpad; m += kWave) {
      float2 e[2], o[2];
#pragma unroll
      for (int src = 0; src < 2; ++src) {
        e[src] = make_float2(0.f, 0.f);
        o[src] = make_float2(0.f, 0.f);
        if (src < nsrc && live && m < T1) {
          const float2* Zs = Zsrc[src];
          float2 zm = Zs[m];
          float2 zc = cconj(Zs[(N - m) % N]);
          float2 xn = make_float2(0.5f * (zm.x + zc.x), 0.5f * (zm.y + zc.y));
          float2 d = csub(zm, zc);
          float2 xs = make_float2(0.5f * d.y, -0.5f * d.x);
          e[src] = make_float2((xn.x + xs.x) * s, (xn.y + xs.y) * s);
          o[src] = make_float2((xn.x - xs.x) * s, (xn.y - xs.y) * s);
        }
      }
      auto put_w = [&](int c, int p, float v) {
        a.job.waist0[waist_index(c, a.job.ldw, a.JHP, m, p, i)] = v;
      };
      if (!a.job.vector) {
        const int c = a.job.col0 + 2 * item;
        put_w(c, 0, e[0].x);
        put_w(c + 1, 0, e[0].y);
        put_w(c, 1, o[0].x);
        put_w(c + 1, 1, o[0].y);
      } else {
        // level columns [iV, iU] = [-V.im, V.re, -U.im, U.re] (fold p); the
        // dP/dphi segment reads [U, -V] = [a3, -a2, -a1, a0] of the same
        // columns through the GEMM's column map.
        const int c = a.job.col0 + 4 * item;
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          const float2 U = p == 0 ? e[0] : o[0];
          const float2 V = p == 0 ? e[1] : o[1];
          const float va[4] = {-V.y, V.x, -U.y, U.x};
#pragma unroll
          for (int cl = 0; cl < 4; ++cl) put_w(c + cl, p, va[cl]);
        }
      }
    }
    __syncthreads();  // buffers are rewritten by the next group
  }
}

// ---------------------------------------------------------------------------
// Synthesis: folded waist -> grid rows.
// ---------------------------------------------------------------------------
// A block walks `gpb` consecutive row groups of one field; each wavefront
// fetches its next row's E/O values (PF per lane and quantity) while it
// transforms the current one.  LDS: the N twiddles, then two N-point buffers
// per wavefront.
template <int PF>
__global__ void __launch_bounds__(kFftThreads) irfft_unfold_kernel(UnfoldArgs a) {
  extern __shared__ float2 lds[];
  const int N = a.fft.N;
  const int T1 = a.T + 1;
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  const int chunk = blockIdx.x % a.nchunks;
  const int f = blockIdx.x / a.nchunks;
  if (f >= a.job.nfields) return;  // block-uniform
  float2* tw = lds;
  float2* b0 = lds + N + (size_t)w * 2 * N;
  float2* b1 = b0 + N;
  for (int k = threadIdx.x; k < N; k += blockDim.x) tw[k] = a.fft.tw[k];
  const int c = a.job.col0 + a.job.fstride * f;
  const float* W = a.job.waist;
  const int g0 = chunk * a.gpb;
  const int g1 = min(g0 + a.gpb, a.ngroups);
  // orders per lane, m = lane + 64 j < T+1 <= N/2: half the row depth
  constexpr int MP = PF > 0 ? PF / 2 : 1;
  float er[MP], ei[MP], orr[MP], oi[MP];
  auto fetch = [&](int g) {
    if constexpr (PF > 0) {
      const int i = g * a.rp + w;
      if (i >= a.JH) return;
#pragma unroll
      for (int j = 0; j < MP; ++j) {
        const int m = lane + kWave * j;
        const bool ok = m < T1;
        er[j] = ok ? W[waist_index(c, a.job.ldw, a.JHP, m, 0, i)] : 0.f;
        ei[j] = ok ? W[waist_index(c + 1, a.job.ldw, a.JHP, m, 0, i)] : 0.f;
        orr[j] = ok ? W[waist_index(c, a.job.ldw, a.JHP, m, 1, i)] : 0.f;
        oi[j] = ok ? W[waist_index(c + 1, a.job.ldw, a.JHP, m, 1, i)] : 0.f;
      }
    }
  };
  // spectrum value Z at order m from E, O
  auto zval = [&](float2 E, float2 O, float s, int m, bool lower) {
    float2 A = make_float2((E.x + O.x) * s, (E.y + O.y) * s);
    float2 B = make_float2((E.x - O.x) * s, (E.y - O.y) * s);
    if (a.job.imul) {
      const float k = a.job.imul == 2 ? (float)m : 1.f;  // i m or i
      A = make_float2(-k * A.y, k * A.x);
      B = make_float2(-k * B.y, k * B.x);
    }
    if (m == 0) {
      A.y = 0.f;
      B.y = 0.f;
    }
    return lower ? make_float2(A.x - B.y, A.y + B.x)     // A + iB
                 : make_float2(A.x + B.y, -A.y + B.x);   // conj(A) + i conj(B)
  };

  fetch(g0);
  for (int g = g0; g < g1; ++g) {
    const int i = g * a.rp + w;
    const bool live = i < a.JH;
    if (live) {
      const float s = a.job.rowscale[i];
      // Z[k] = A + iB at k = m <= N/2, conj(A) + i conj(B) at k = N - m; zero above T
      for (int k = lane; k < N; k += kWave) b0[k] = make_float2(0.f, 0.f);
      if constexpr (PF > 0) {
#pragma unroll
        for (int j = 0; j < MP; ++j) {
          const int m = lane + kWave * j;
          if (m < T1) {
            const float2 E = make_float2(er[j], ei[j]), O = make_float2(orr[j], oi[j]);
            b0[m] = zval(E, O, s, m, true);
            if (m > 0) b0[N - m] = zval(E, O, s, m, false);
          }
        }
      } else {
        for (int m = lane; m < T1; m += kWave) {
          const float2 E = make_float2(W[waist_index(c, a.job.ldw, a.JHP, m, 0, i)],
                                       W[waist_index(c + 1, a.job.ldw, a.JHP, m, 0, i)]);
          const float2 O = make_float2(W[waist_index(c, a.job.ldw, a.JHP, m, 1, i)],
                                       W[waist_index(c + 1, a.job.ldw, a.JHP, m, 1, i)]);
          b0[m] = zval(E, O, s, m, true);
          if (m > 0) b0[N - m] = zval(E, O, s, m, false);
        }
      }
    }
    if (g + 1 < g1) fetch(g + 1);  // next row's E/O load during this FFT
    __syncthreads();
    float2* x = wave_fft(b0, b1, a.fft, tw, +1, lane);
    if (live) {
      const size_t plane = (size_t)a.nlat * N;
      float* rn = a.job.grid + (size_t)f * plane + (size_t)(a.nlat - 1 - i) * N;
      float* rs = a.job.grid + (size_t)f * plane + (size_t)i * N;
      for (int k = lane; k < N; k += kWave) {
        float2 v = x[k];
        rn[k] = a.job.post_exp ? expf(v.x) : v.x;
        rs[k] = a.job.post_exp ? expf(v.y) : v.y;
      }
    }
    __syncthreads();  // buffers are rewritten by the next group
  }
}

// ---------------------------------------------------------------------------
// Wave-resident FFTs of N = 64 L points (L = 2, 3, 4, 6, 8, 12, 16, 24)
// ---------------------------------------------------------------------------
// One wavefront transforms F sequences on its own, with no block barrier:
// with n = 64 n1 + n2 and k = k1 + L (c + 8 d), n2 = 8 a + b,
//   A (registers): lane n2 holds x[64 n1 + n2], runs the L-point DFT over n1
//                  and multiplies by w_N^(n2 k1);
//   B (LDS transpose): the 8-point DFT over a of each (k1, b), times
//                  w_64^(b c);
//   C (LDS transpose): the 8-point DFT over b of each (k1, c), giving
//                  X[k1 + L c + 8 L d], d = 0..7, stored in natural order.
// The small DFTs are unrolled at compile time (mixed radix 4/3/2, constant
// twiddles); the w_N and w_64 twiddles come from the block's copy of the FP64
// table.  Each wavefront owns an LDS span of L rows of 72 points.

constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double taylor_cos(double x) {
  double t = 1.0, s = 1.0;
  for (int n = 1; n < 40; ++n) {
    t *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
    s += t;
  }
  return s;
}
constexpr double taylor_sin(double x) {
  double t = x, s = x;
  for (int n = 1; n < 40; ++n) {
    t *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
    s += t;
  }
  return s;
}
// cos and sin of 2 pi j / 48, rounded once to fp32 (exact zeros snapped);
// Ring48<double> keeps them in FP64 for the FP64 FFTs
template <class R>
struct Ring48 {
  R c[48];
  R s[48];
};
template <class R>
constexpr Ring48<R> make_ring48() {
  Ring48<R> r{};
  for (int j = 0; j < 48; ++j) {
    const double x = kTwoPi * j / 48.0;
    const double c = taylor_cos(x), sn = taylor_sin(x);
    r.c[j] = (R)((c < 1e-12 && c > -1e-12) ? 0.0 : c);
    r.s[j] = (R)((sn < 1e-12 && sn > -1e-12) ? 0.0 : sn);
  }
  return r;
}
constexpr Ring48<float> kRing48 = make_ring48<float>();
constexpr Ring48<double> kRing48d = make_ring48<double>();

template <class Fn, int... I>
__device__ __forceinline__ void static_for_impl(Fn&& fn, std::integer_sequence<int, I...>) {
  (fn(std::integral_constant<int, I>{}), ...);
}
// fn(integral_constant<int, i>) for i = 0 .. N-1, unrolled at compile time
template <int N, class Fn>
__device__ __forceinline__ void static_for(Fn&& fn) {
  static_for_impl(fn, std::make_integer_sequence<int, N>{});
}

// x * e^(S 2 pi i E / R), R | 48; C is float2 or double2
template <int E, int R, int S, class C>
__device__ __forceinline__ C rot(C x) {
  constexpr int e = ((E % R) + R) % R;
  if constexpr (e == 0) {
    return x;
  } else if constexpr (2 * e == R) {
    return cmk(-x.x, -x.y);
  } else if constexpr (4 * e == R) {  // S i
    return S > 0 ? cmk(-x.y, x.x) : cmk(x.y, -x.x);
  } else if constexpr (4 * e == 3 * R) {  // -S i
    return S > 0 ? cmk(x.y, -x.x) : cmk(-x.y, x.x);
  } else if constexpr (std::is_same<C, float2>::value) {
    constexpr int j = e * (48 / R);
    constexpr float c = kRing48.c[j], sn = S * kRing48.s[j];
    return make_float2(x.x * c - x.y * sn, x.x * sn + x.y * c);
  } else {
    constexpr int j = e * (48 / R);
    constexpr double c = kRing48d.c[j], sn = S * kRing48d.s[j];
    return make_double2(x.x * c - x.y * sn, x.x * sn + x.y * c);
  }
}

// In-register DFT of R points, natural order in and out, sign S.
template <int R, int S, class C>
__device__ __forceinline__ void dft(C (&v)[R]) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
    const C a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 3) {
    const C t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    C m1, m2;
    if constexpr (std::is_same<C, float2>::value) {
      constexpr float h = 0.86602540378443864676f * S;  // sin(S 2 pi / 3)
      m1 = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
      m2 = make_float2(-h * t2.y, h * t2.x);
    } else {
      constexpr double h = 0.86602540378443864676 * S;
      m1 = make_double2(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
      m2 = make_double2(-h * t2.y, h * t2.x);
    }
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m1, m2);
    v[2] = csub(m1, m2);
  } else if constexpr (R == 4) {
    const C a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    const C a2 = cadd(v[1], v[3]), a3 = rot<1, 4, S>(csub(v[1], v[3]));
    v[0] = cadd(a0, a2);
    v[2] = csub(a0, a2);
    v[1] = cadd(a1, a3);
    v[3] = csub(a1, a3);
  } else {
    // n = Q n1 + n2, k = k1 + P k2:
    // X[k1 + P k2] = sum_n2 w_Q^(n2 k2) w_R^(n2 k1) sum_n1 x[Q n1 + n2] w_P^(n1 k1)
    constexpr int P = (R % 4 == 0) ? 4 : ((R % 3 == 0) ? 3 : 2);
    constexpr int Q = R / P;
    C y[R];
    static_for<Q>([&](auto n2_) {
      constexpr int n2 = decltype(n2_)::value;
      C t[P];
      static_for<P>([&](auto n1_) { t[n1_] = v[Q * decltype(n1_)::value + n2]; });
      dft<P, S>(t);
      static_for<P>([&](auto k1_) {
        constexpr int k1 = decltype(k1_)::value;
        y[n2 * P + k1] = rot<n2 * k1, R, S>(t[k1]);
      });
    });
    static_for<P>([&](auto k1_) {
      constexpr int k1 = decltype(k1_)::value;
      C t[Q];
      static_for<Q>([&](auto n2_) { t[n2_] = y[decltype(n2_)::value * P + k1]; });
      dft<Q, S>(t);
      static_for<Q>([&](auto k2_) { v[k1 + P * decltype(k2_)::value] = t[k2_]; });
    });
  }
}

// Order this wavefront's LDS accesses (LDS operations of one wavefront
// complete in issue order once lgkmcnt drains).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kRowSpan = 72;  // LDS row of the transposes: 64 points + 8 pad
template <int L>
constexpr int wave_span() { return L * kRowSpan; }  // float2 per sequence (>= 64 L)
// Between the wavefronts' spans: 4 float2 (8 banks) of padding, so that the
// eight wavefronts' copies of one 16-byte piece (read or written by eight
// consecutive lanes) fall in eight different bank groups.
constexpr int kSpanPad = 4;

// Twiddle entries of the wave-resident kernels' LDS table: N + 64.
template <int L>
constexpr int wave_tw() { return 64 * L + 64; }

// The wave-resident FFT's twiddles in LDS, in the order its lanes read them:
// [k1][lane] = w^{lane k1} for the first pass (k1 < L), then [c][b] =
// w^{b c L} for the second (c, b < 8), w = e^{-2 pi i / N} from the table.
// Consecutive lanes read consecutive entries (and in the second pass the
// eight b-groups of a wavefront read the same eight), so neither pass has
// bank conflicts; read at w^{lane k1} and w^{b c L} straight from the table,
// 16-byte entries collided up to 8 ways.
template <int L, class C>
__device__ __forceinline__ void fill_wave_twiddles(C* tw, const C* table) {
  constexpr int N = 64 * L;
  for (int k = threadIdx.x; k < wave_tw<L>(); k += blockDim.x)
    tw[k] = k < N ? table[(k >> 6) * (k & 63)] : table[((k - N) & 7) * ((k - N) >> 3) * L];
}

// F sequences of N = 64 L points: v[f][n1] = x_f[64 n1 + lane] on entry; on
// return buf[f * span + k] holds X_f[k], k < N.  tw: fill_wave_twiddles.
template <int L, int F, int S, class C>
__device__ __forceinline__ void wave_fft64(C (&v)[F][L], C* buf, const C* tw, int lane) {
  constexpr int SP = wave_span<L>();
  static_for<F>([&](auto f_) {
    constexpr int f = decltype(f_)::value;
    dft<L, S>(v[f]);
    static_for<L>([&](auto k1_) {
      constexpr int k1 = decltype(k1_)::value;
      C x = v[f][k1];
      if constexpr (k1 > 0) x = cmul(x, twiddle(tw, 64 * k1 + lane, S));
      buf[f * SP + k1 * kRowSpan + lane] = x;
    });
  });
  wave_sync();
  constexpr int U = F * L * 8, RND = (U + 63) / 64;
  C w[RND][8];
  // B: units (f, k1, b)
  static_for<RND>([&](auto r_) {
    constexpr int r = decltype(r_)::value;
    const int u = lane + 64 * r;
    if (u < U) {
      const int f = u / (8 * L), k1 = (u >> 3) % L, b = u & 7;
      const C* row = buf + f * SP + k1 * kRowSpan + b;
      static_for<8>([&](auto a_) { w[r][a_] = row[8 * decltype(a_)::value]; });
      dft<8, S>(w[r]);
      static_for<8>([&](auto c_) {
        constexpr int c = decltype(c_)::value;
        if constexpr (c > 0) w[r][c] = cmul(w[r][c], twiddle(tw, 64 * L + 8 * c + b, S));
      });
    }
  });
  wave_sync();
  // S2[k1][c][b ^ c]: the B writes (lanes over b) and the C reads (lanes
  // over c) both hit distinct banks (measured)
  static_for<RND>([&](auto r_) {
    constexpr int r = decltype(r_)::value;
    const int u = lane + 64 * r;
    if (u < U) {
      const int f = u / (8 * L), k1 = (u >> 3) % L, b = u & 7;
      C* row = buf + f * SP + k1 * kRowSpan;
      static_for<8>([&](auto c_) {
        constexpr int c = decltype(c_)::value;
        row[8 * c + (b ^ c)] = w[r][c];
      });
    }
  });
  wave_sync();
  // C: units (f, k1, c)
  static_for<RND>([&](auto r_) {
    constexpr int r = decltype(r_)::value;
    const int u = lane + 64 * r;
    if (u < U) {
      const int f = u / (8 * L), k1 = (u >> 3) % L, c = u & 7;
      const C* row = buf + f * SP + k1 * kRowSpan + 8 * c;
      static_for<8>([&](auto b_) { w[r][b_] = row[decltype(b_)::value ^ c]; });
      dft<8, S>(w[r]);
    }
  });
  wave_sync();
  static_for<RND>([&](auto r_) {
    constexpr int r = decltype(r_)::value;
    const int u = lane + 64 * r;
    if (u < U) {
      const int f = u / (8 * L), k1 = (u >> 3) % L, c = u & 7;
      C* out = buf + f * SP + k1 + L * c;
      static_for<8>([&](auto d_) { out[8 * L * decltype(d_)::value] = w[r][d_]; });
    }
  });
  wave_sync();
}

// Blocks per CU that a wave-resident launch's LDS allows (2 when a block
// takes at most half of the 160 KB).  The scalar fold and the unfold are
// compiled for that many (their second launch bound, waves per SIMD, is
// blocks x RB / 4), so that registers do not cap their residency below what
// LDS allows; the vector fold's LDS allows one block.
template <int L, int F, int RB>
constexpr int wave_blocks_per_cu() {
  return (wave_tw<L>() + RB * (F * L * kRowSpan + kSpanPad)) * 8 * 2 <= 160 * 1024 ? 2 : 1;
}
template <int L, int F, int RB>
constexpr int fold_waves_per_simd() {
  return (F == 1 ? wave_blocks_per_cu<L, 1, RB>() : 1) * RB / 4;
}

// Wavefronts (rows) per block of the wave-resident kernels, RB: with 8 a
// (column, m-tile, parity) run of 8 rows x 4 orders is one whole 128-byte
// line, so no line of the waist is shared between blocks; 4 where 8 rows'
// LDS does not fit (the longest rows).

// Analysis: one wavefront per (field, folded row); F = 2 transforms u and v
// of a vector level together.  A group is kWaveBlock consecutive rows of one
// field (JHP is a multiple of 32): each wavefront folds its row's spectrum
// into registers and parks the values in its LDS span, then the block writes
// every (column, m-tile, parity) run of RB rows x 4 orders as one piece.  Blocks are persistent: while a group is transformed, the rows of
// the block's next group are already being loaded into registers.
template <int L, int F, int kWaveBlock>
__global__ void __launch_bounds__(64 * kWaveBlock, (fold_waves_per_simd<L, F, kWaveBlock>()))
    rfft_fold_wave_kernel(FoldArgs a) {
  extern __shared__ float2 lds[];
  constexpr int N = 64 * L, SP = wave_span<L>();
  constexpr int NV = 4 * F;  // folded values per order: columns x parities
  static_assert(kMT == 4, "a row's four orders are one 16-byte vector");
  typedef __attribute__((ext_vector_type(4))) float f4;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  constexpr int WS = F * SP + kSpanPad;  // span stride per wavefront
  float2* tw = lds;
  float2* buf = lds + wave_tw<L>() + (size_t)w * WS;
  const int ngroups = a.job.nfields * (a.JHP / kWaveBlock);
  const int T1 = a.T + 1;
  const int MT = (T1 + kMT - 1) / kMT;  // m-tiles
  const int MO = MT * kMT;               // orders incl. the last tile's padding
  const size_t plane = (size_t)a.nlat * N;
  int g = blockIdx.x;
  if (g >= ngroups) return;  // block-uniform
  float2 v[F][L];
  auto load_rows = [&](int grp) {
    const int item = grp * kWaveBlock + w;
    const int fld = item / a.JHP, i = item % a.JHP;
    if (i >= a.JH) return;
    static_for<F>([&](auto f_) {
      constexpr int f = decltype(f_)::value;
      const float* gp = (f == 0 ? a.job.grid0 : a.job.grid1) + (size_t)fld * plane;
      const float* rn = gp + (size_t)(a.nlat - 1 - i) * N;
      const float* rs = gp + (size_t)i * N;
      // a scalar field's grid offset (x - 0 = x exactly)
      const float off = (F == 1 && a.job.goff) ? a.job.goff[fld] : 0.f;
      static_for<L>([&](auto n1_) {
        constexpr int n1 = decltype(n1_)::value;
        v[f][n1] = make_float2(rn[64 * n1 + lane] - off, rs[64 * n1 + lane] - off);
      });
    });
  };
  load_rows(g);
  fill_wave_twiddles<L>(tw, a.fft.tw);
  __syncthreads();
  for (; g < ngroups; g += gridDim.x) {
    const int item0 = g * kWaveBlock;
    const int fld = item0 / a.JHP, i0 = item0 % a.JHP, i = i0 + w;
    const bool live = i < a.JH;
    if (live) wave_fft64<L, F, -1>(v, buf, tw, lane);  // v is free from here on
    if (g + (int)gridDim.x < ngroups) load_rows(g + gridDim.x);
    // fold into registers, one order per lane and round (consecutive orders
    // across lanes, so the spectrum reads are contiguous): val[r][e] for
    // order m = lane + 64 r, e = 2 * column + parity
    const float s0 = live ? a.job.rowscale[i] : 0.f;
    const float s1 = live && a.job.rowscale_hi ? a.job.rowscale_hi[i] : s0;
    constexpr int RO = (N / 2 + 4 + 63) / 64;  // rounds: MO <= N/2 + 4
    float val[RO][NV];
#pragma unroll
    for (int r = 0; r < RO; ++r) {
      const int m = lane + 64 * r;
      const float s = m < a.job.m_hi ? s0 : s1;
      float2 e[2], o[2];
#pragma unroll
      for (int f = 0; f < 2; ++f) {
        e[f] = make_float2(0.f, 0.f);
        o[f] = make_float2(0.f, 0.f);
        if (f < F && live && m < T1) {
          const float2 zm = buf[f * SP + m];
          const float2 zc = cconj(buf[f * SP + (N - m) % N]);
          const float2 xn = make_float2(0.5f * (zm.x + zc.x), 0.5f * (zm.y + zc.y));
          const float2 d = csub(zm, zc);
          const float2 xs = make_float2(0.5f * d.y, -0.5f * d.x);
          e[f] = make_float2((xn.x + xs.x) * s, (xn.y + xs.y) * s);
          o[f] = make_float2((xn.x - xs.x) * s, (xn.y - xs.y) * s);
        }
      }
      if constexpr (F == 1) {
        val[r][0] = e[0].x;
        val[r][1] = o[0].x;
        val[r][2] = e[0].y;
        val[r][3] = o[0].y;
      } else {
        // level columns [iV, iU] = [-V.im, V.re, -U.im, U.re]
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          const float2 U = p == 0 ? e[0] : o[0];
          const float2 V = p == 0 ? e[1] : o[1];
          val[r][0 + p] = -V.y;
          val[r][2 + p] = V.x;
          val[r][4 + p] = -U.y;
          val[r][6 + p] = U.x;
        }
      }
    }
    wave_sync();  // the spectrum is read; the span now holds stage[e][m]
    float* stage = reinterpret_cast<float*>(buf);
#pragma unroll
    for (int r = 0; r < RO; ++r) {
      const int m = lane + 64 * r;
      if (m < MO) {
#pragma unroll
        for (int e = 0; e < NV; ++e) stage[(size_t)e * MO + m] = val[r][e];
      }
    }
    __syncthreads();
    // block write: run (e, mt) = rows i0 .. i0+RB-1 x orders 4 mt .. 4 mt + 3,
    // one contiguous piece; consecutive lanes take its rows (16 bytes each).
    // e = rn / MT is floor((rn + 1/2) * (1/MT)) in fp32, exact for rn < 2^12
    // and MT <= 256: the product is within 2^-11 of (rn + 1/2) / MT, which is
    // at least 1 / (2 MT) >= 2^-9 from the nearest integer.
    const int ncolf = 2 * F;  // columns per field or level
    const float inv_mt = 1.0f / (float)MT;
    static_assert(NV * (N / 8 + 1) < 4096 && N / 8 + 1 <= 256, "range of the fp32 quotient");
    for (int u = threadIdx.x; u < NV * MT * kWaveBlock; u += blockDim.x) {
      const int ww = u % kWaveBlock, rn = u / kWaveBlock;
      const int e = (int)(((float)rn + 0.5f) * inv_mt), mt = rn - e * MT;
      const int c = a.job.col0 + ncolf * fld + (e >> 1), p = e & 1;
      *reinterpret_cast<f4*>(a.job.waist0 + waist_index(c, a.job.ldw, a.JHP, mt * kMT, p, i0 + ww)) =
          *reinterpret_cast<const f4*>(reinterpret_cast<const float*>(lds + wave_tw<L>() + (size_t)ww * WS) +
                                       ((size_t)e * MT + mt) * 4);
    }
    __syncthreads();  // the spans are free for the next group
  }
}

// Synthesis: one wavefront per (field, folded row).  A group is kWaveBlock
// consecutive rows of one field: the block reads every (column, m-tile,
// parity) run of RB rows x 4 orders of the waist as one piece and hands each
// row its values through LDS.  Blocks are persistent:
// the next group's waist runs are loaded into registers while the current
// group is transformed.
template <int L, int kWaveBlock>
__global__ void __launch_bounds__(64 * kWaveBlock,
                                  (wave_blocks_per_cu<L, 1, kWaveBlock>() * kWaveBlock / 4))
    irfft_unfold_wave_kernel(UnfoldArgs a) {
  extern __shared__ float2 lds[];
  constexpr int N = 64 * L, SP = wave_span<L>();
  static_assert(kMT == 4, "a row's four orders are one 16-byte vector");
  typedef __attribute__((ext_vector_type(4))) float f4;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  constexpr int WS = SP + kSpanPad;  // span stride per wavefront
  float2* tw = lds;
  float2* buf = lds + wave_tw<L>() + (size_t)w * WS;
  const int JH4 = (a.JH + kWaveBlock - 1) / kWaveBlock;  // row groups
  const int ngroups = a.job.nfields * JH4;
  const int T1 = a.T + 1;
  const int MT = (T1 + kMT - 1) / kMT;
  const int MO = MT * kMT;
  const float* W = a.job.waist;
  const size_t plane = (size_t)a.nlat * N;
  int g = blockIdx.x;
  if (g >= ngroups) return;  // block-uniform
  // a thread's share of one group's runs: stage[e][mt][z], e = 2 * column +
  // parity; consecutive lanes take one run's rows (16 bytes each).  Where a
  // run sits relative to the group's first row, in the waist and in LDS,
  // is the same for every group, so it is worked out once:
  //   waist  W + c ldw + 4 i0 + roff[r],  roff = (e >> 1) ldw + 4 ((2 mt + (e & 1)) JHP + ww)
  //   LDS    span ww, float 4 (e MT + mt)
  // (the waist layout W[c][m/4][p][i][m%4]); rvalid bit r marks r < 4 MT.
  constexpr int RU = (4 * (N / 8 + 1) + 63) / 64;  // 4 MT <= N/2 + 4 runs
  f4 run[RU];
  unsigned roff[RU], loff[RU], rvalid = 0;
#pragma unroll
  for (int r = 0; r < RU; ++r) {
    const int u = threadIdx.x + 64 * kWaveBlock * r;
    const int ww = u % kWaveBlock, rn = u / kWaveBlock;
    const bool ok = rn < 4 * MT;
    const unsigned e = ok ? (unsigned)(rn / MT) : 0u, mt = ok ? (unsigned)(rn % MT) : 0u;
    rvalid |= (ok ? 1u : 0u) << r;
    roff[r] = (e >> 1) * (unsigned)a.job.ldw + 4u * ((2u * mt + (e & 1u)) * (unsigned)a.JHP + ww);
    loff[r] = (unsigned)(2 * wave_tw<L>() + 2 * ww * WS) + 4u * (e * (unsigned)MT + mt);
  }
  auto load_runs = [&](int grp) {
    const int fld = grp / JH4, i0 = (grp % JH4) * kWaveBlock;
    const float* base = W + (size_t)(a.job.col0 + a.job.fstride * fld) * a.job.ldw + 4 * i0;
#pragma unroll
    for (int r = 0; r < RU; ++r)
      if (rvalid & (1u << r)) run[r] = *reinterpret_cast<const f4*>(base + roff[r]);
  };
  load_runs(g);
  fill_wave_twiddles<L>(tw, a.fft.tw);
  for (; g < ngroups; g += gridDim.x) {
    // the lane index is made opaque in each group, so that the FFT's
    // lane-derived LDS addresses are recomputed there rather than held in
    // registers across the loop: the kernel then fits the 128 VGPRs of two
    // blocks per CU without spilling
    int lane = threadIdx.x & 63;
    asm volatile("" : "+v"(lane));
    const int fld = g / JH4, i0 = (g % JH4) * kWaveBlock, i = i0 + w;
    const bool live = i < a.JH;
#pragma unroll
    for (int r = 0; r < RU; ++r)
      if (rvalid & (1u << r)) *reinterpret_cast<f4*>(reinterpret_cast<float*>(lds) + loff[r]) = run[r];
    __syncthreads();  // spans (and, the first time, the twiddles) are in LDS
    if (g + (int)gridDim.x < ngroups) load_runs(g + gridDim.x);
    if (live) {
      // one order per lane and round: stage[e][m], e = 2 * column + parity
      constexpr int RO = (N / 2 + 4 + 63) / 64;  // rounds: MO <= N/2 + 4
      float er[RO], ei[RO], orr[RO], oi[RO];
      const float* stage = reinterpret_cast<const float*>(buf);
#pragma unroll
      for (int r = 0; r < RO; ++r) {
        const int m = lane + 64 * r;
        const bool in = m < MO;
        er[r] = in ? stage[m] : 0.f;
        orr[r] = in ? stage[(size_t)MO + m] : 0.f;
        ei[r] = in ? stage[(size_t)2 * MO + m] : 0.f;
        oi[r] = in ? stage[(size_t)3 * MO + m] : 0.f;
      }
      wave_sync();  // the staged values are in registers; the span becomes the spectrum
      const float s0 = a.job.rowscale[i];
      const float s1 = a.job.rowscale_hi ? a.job.rowscale_hi[i] : s0;
      // Z[k] = A + iB at k = m, conj(A) + i conj(B) at k = N - m; zero above T
      for (int k = lane; k < N; k += 64) buf[k] = make_float2(0.f, 0.f);
      wave_sync();
#pragma unroll
      for (int r = 0; r < RO; ++r) {
        const int m = lane + 64 * r;
        if (m >= T1) continue;
        const float s = m < a.job.m_hi ? s0 : s1;
        float2 A = make_float2((er[r] + orr[r]) * s, (ei[r] + oi[r]) * s);
        float2 B = make_float2((er[r] - orr[r]) * s, (ei[r] - oi[r]) * s);
        if (a.job.imul) {
          const float k = a.job.imul == 2 ? (float)m : 1.f;  // i m or i
          A = make_float2(-k * A.y, k * A.x);
          B = make_float2(-k * B.y, k * B.x);
        }
        if (m == 0) {
          A.y = 0.f;
          B.y = 0.f;
        }
        buf[m] = make_float2(A.x - B.y, A.y + B.x);
        if (m > 0) buf[N - m] = make_float2(A.x + B.y, -A.y + B.x);
      }
      wave_sync();
      float2 v[1][L];
      static_for<L>([&](auto n1_) { v[0][n1_] = buf[64 * decltype(n1_)::value + lane]; });
      wave_sync();
      wave_fft64<L, 1, +1>(v, buf, tw, lane);
      float* rn = a.job.grid + (size_t)fld * plane + (size_t)(a.nlat - 1 - i) * N;
      float* rs = a.job.grid + (size_t)fld * plane + (size_t)i * N;
      const bool post_exp = a.job.post_exp != 0;
      static_for<L>([&](auto q_) {
        const int k = lane + 64 * decltype(q_)::value;
        const float2 x = buf[k];
        rn[k] = post_exp ? expf(x.x) : x.x;
        rs[k] = post_exp ? expf(x.y) : x.y;
      });
    }
    __syncthreads();  // every span is consumed before the next group's runs land
  }
}

// ---------------------------------------------------------------------------
// FP64 wave-resident FFTs of the vector operators (u, v) for their FP64
// correction.  The large low orders of a wind field leak fp32 rounding into
// every order of an fp32 FFT; relative to the small high-degree vorticity and
// divergence those leaks exceed 1e-6.  Here one wavefront transforms one
// field's mirror row pair in FP64 (grid values fp32 as stored), and the
// folded values go to the fp32 waist (orders the split-BF16 GEMM contracts)
// and, for m < m0, in FP64 to the low waist (orders the FP64 low-order kernels
// contract).  The two parts take their own row scales (rowscale64 for the
// low orders, rowscale64_hi for the rest: the recurrence form's extra
// 1/cos(phi) applies to the GEMM's orders only).  Same row grouping,
// persistence and block writes as the fp32 kernels above; LDS spans hold
// double2.
// ---------------------------------------------------------------------------
template <int L, int kWaveBlock>
__global__ void __launch_bounds__(64 * kWaveBlock) rfft_fold_wave64_kernel(FoldArgs a) {
  extern __shared__ double2 lds_d[];
  constexpr int N = 64 * L, SP = wave_span<L>();
  constexpr int NV = 4;  // folded values per order: 2 columns x 2 parities
  typedef __attribute__((ext_vector_type(4))) float f4;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  constexpr int WS = SP + kSpanPad;
  double2* tw = lds_d;
  double2* buf = lds_d + wave_tw<L>() + (size_t)w * WS;
  const int ngroups = a.job.nfields * (a.JHP / kWaveBlock);
  const int T1 = a.T + 1;
  const int MT = (T1 + kMT - 1) / kMT;
  const int MO = MT * kMT;
  const size_t plane = (size_t)a.nlat * N;
  int g = blockIdx.x;
  if (g >= ngroups) return;  // block-uniform
  float2 pre[L];             // the next mirror pair as stored (north, south)
  auto load_rows = [&](int grp) {
    const int item = grp * kWaveBlock + w;
    const int fld = item / a.JHP, i = item % a.JHP;
    if (i >= a.JH) return;
    const float* gp = a.job.grid0 + (size_t)fld * plane;
    const float* rn = gp + (size_t)(a.nlat - 1 - i) * N;
    const float* rs = gp + (size_t)i * N;
    static_for<L>([&](auto n1_) {
      constexpr int n1 = decltype(n1_)::value;
      pre[n1] = make_float2(rn[64 * n1 + lane], rs[64 * n1 + lane]);
    });
  };
  load_rows(g);
  fill_wave_twiddles<L>(tw, a.fft.tw64);
  __syncthreads();
  for (; g < ngroups; g += gridDim.x) {
    const int item0 = g * kWaveBlock;
    const int fld = item0 / a.JHP, i0 = item0 % a.JHP, i = i0 + w;
    const bool live = i < a.JH;
    if (live) {
      double2 v[1][L];
      static_for<L>([&](auto n1_) {
        constexpr int n1 = decltype(n1_)::value;
        v[0][n1] = make_double2((double)pre[n1].x, (double)pre[n1].y);
      });
      wave_fft64<L, 1, -1>(v, buf, tw, lane);
    }
    if (g + (int)gridDim.x < ngroups) load_rows(g + gridDim.x);
    const double s = live ? a.job.rowscale64[i] : 0.0;
    const double sh = live && a.job.rowscale64_hi ? a.job.rowscale64_hi[i] : s;
    double2* low = a.job.low + (size_t)fld * a.job.low_stride;
    constexpr int RO = (N / 2 + 4 + 63) / 64;  // rounds: MO <= N/2 + 4
    float val[RO][NV];
#pragma unroll
    for (int r = 0; r < RO; ++r) {
      const int m = lane + 64 * r;
      double2 e = make_double2(0.0, 0.0), o = make_double2(0.0, 0.0);
      if (live && m < T1) {
        const double2 zm = buf[m];
        const double2 zc = cconj(buf[(N - m) % N]);
        const double2 xn = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
        const double2 d = csub(zm, zc);
        const double2 xs = make_double2(0.5 * d.y, -0.5 * d.x);
        const double sc = m < a.job.m0 ? s : sh;
        e = make_double2((xn.x + xs.x) * sc, (xn.y + xs.y) * sc);
        o = make_double2((xn.x - xs.x) * sc, (xn.y - xs.y) * sc);
        if (m < a.job.m0) {
          low[((size_t)m * 2 + 0) * a.JHP + i] = e;
          low[((size_t)m * 2 + 1) * a.JHP + i] = o;
        }
      }
      // i X: columns [-X.im, X.re], e = 2 * column + parity
      val[r][0] = (float)-e.y;
      val[r][1] = (float)-o.y;
      val[r][2] = (float)e.x;
      val[r][3] = (float)o.x;
    }
    wave_sync();  // the spectrum is read; the span now holds stage[e][m]
    float* stage = reinterpret_cast<float*>(buf);
#pragma unroll
    for (int r = 0; r < RO; ++r) {
      const int m = lane + 64 * r;
      if (m < MO) {
#pragma unroll
        for (int e = 0; e < NV; ++e) stage[(size_t)e * MO + m] = val[r][e];
      }
    }
    __syncthreads();
    // block write as in rfft_fold_wave_kernel (e = rn / MT from the fp32
    // reciprocal, exact in this range)
    const float inv_mt = 1.0f / (float)MT;
    static_assert(NV * (N / 8 + 1) < 4096 && N / 8 + 1 <= 256, "range of the fp32 quotient");
    for (int u = threadIdx.x; u < NV * MT * kWaveBlock; u += blockDim.x) {
      const int ww = u % kWaveBlock, rn = u / kWaveBlock;
      const int e = (int)(((float)rn + 0.5f) * inv_mt), mt = rn - e * MT;
      const int c = a.job.col0 + a.job.fstride * fld + (e >> 1), p = e & 1;
      *reinterpret_cast<f4*>(a.job.waist0 + waist_index(c, a.job.ldw, a.JHP, mt * kMT, p, i0 + ww)) =
          *reinterpret_cast<const f4*>(reinterpret_cast<const float*>(lds_d + wave_tw<L>() + (size_t)ww * WS) +
                                       ((size_t)e * MT + mt) * 4);
    }
    __syncthreads();  // the spans are free for the next group
  }
}

template <int L, int kWaveBlock>
__global__ void __launch_bounds__(64 * kWaveBlock) irfft_unfold_wave64_kernel(UnfoldArgs a) {
  extern __shared__ double2 lds_d[];
  constexpr int N = 64 * L, SP = wave_span<L>();
  static_assert(kMT == 4, "a row's four orders are one 16-byte vector");
  typedef __attribute__((ext_vector_type(4))) float f4;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  constexpr int WS = SP + kSpanPad;  // span stride per wavefront, double2
  double2* tw = lds_d;
  double2* buf = lds_d + wave_tw<L>() + (size_t)w * WS;
  const int JH4 = (a.JH + kWaveBlock - 1) / kWaveBlock;
  const int ngroups = a.job.nfields * JH4;
  const int T1 = a.T + 1;
  const int MT = (T1 + kMT - 1) / kMT;
  const int MO = MT * kMT;
  const float* W = a.job.waist;
  const size_t plane = (size_t)a.nlat * N;
  int g = blockIdx.x;
  if (g >= ngroups) return;  // block-uniform
  // the fp32 waist runs, staged as in irfft_unfold_wave_kernel (LDS offsets
  // in floats: the twiddles take 4 N, a span 4 WS)
  constexpr int RU = (4 * (N / 8 + 1) + 63) / 64;
  f4 run[RU];
  unsigned roff[RU], loff[RU], rvalid = 0;
#pragma unroll
  for (int r = 0; r < RU; ++r) {
    const int u = threadIdx.x + 64 * kWaveBlock * r;
    const int ww = u % kWaveBlock, rn = u / kWaveBlock;
    const bool ok = rn < 4 * MT;
    const unsigned e = ok ? (unsigned)(rn / MT) : 0u, mt = ok ? (unsigned)(rn % MT) : 0u;
    rvalid |= (ok ? 1u : 0u) << r;
    roff[r] = (e >> 1) * (unsigned)a.job.ldw + 4u * ((2u * mt + (e & 1u)) * (unsigned)a.JHP + ww);
    loff[r] = (unsigned)(4 * wave_tw<L>() + 4 * ww * WS) + 4u * (e * (unsigned)MT + mt);
  }
  auto load_runs = [&](int grp) {
    const int fld = grp / JH4, i0 = (grp % JH4) * kWaveBlock;
    const float* base = W + (size_t)(a.job.col0 + a.job.fstride * fld) * a.job.ldw + 4 * i0;
#pragma unroll
    for (int r = 0; r < RU; ++r)
      if (rvalid & (1u << r)) run[r] = *reinterpret_cast<const f4*>(base + roff[r]);
  };
  load_runs(g);
  fill_wave_twiddles<L>(tw, a.fft.tw64);
  for (; g < ngroups; g += gridDim.x) {
    const int fld = g / JH4, i0 = (g % JH4) * kWaveBlock, i = i0 + w;
    const bool live = i < a.JH;
#pragma unroll
    for (int r = 0; r < RU; ++r)
      if (rvalid & (1u << r)) *reinterpret_cast<f4*>(reinterpret_cast<float*>(lds_d) + loff[r]) = run[r];
    __syncthreads();  // spans (and, the first time, the twiddles) are in LDS
    if (g + (int)gridDim.x < ngroups) load_runs(g + gridDim.x);
    if (live) {
      constexpr int RO = (N / 2 + 4 + 63) / 64;
      float er[RO], ei[RO], orr[RO], oi[RO];
      const float* stage = reinterpret_cast<const float*>(buf);
#pragma unroll
      for (int r = 0; r < RO; ++r) {
        const int m = lane + 64 * r;
        const bool in = m < MO;
        er[r] = in ? stage[m] : 0.f;
        orr[r] = in ? stage[(size_t)MO + m] : 0.f;
        ei[r] = in ? stage[(size_t)2 * MO + m] : 0.f;
        oi[r] = in ? stage[(size_t)3 * MO + m] : 0.f;
      }
      wave_sync();  // the staged values are in registers; the span becomes the spectrum
      const double s = a.job.rowscale64[i];
      const double sh = a.job.rowscale64_hi ? a.job.rowscale64_hi[i] : s;
      const double2* low = a.job.low + (size_t)fld * a.job.low_stride;
      for (int k = lane; k < N; k += 64) buf[k] = make_double2(0.0, 0.0);
      wave_sync();
#pragma unroll
      for (int r = 0; r < RO; ++r) {
        const int m = lane + 64 * r;
        if (m >= T1) continue;
        double2 E = make_double2((double)er[r], (double)ei[r]);
        double2 O = make_double2((double)orr[r], (double)oi[r]);
        double sc = sh;
        if (m < a.job.m0) {
          E = low[((size_t)m * 2 + 0) * a.JHP + i];
          O = low[((size_t)m * 2 + 1) * a.JHP + i];
          sc = s;
        }
        double2 A = make_double2((E.x + O.x) * sc, (E.y + O.y) * sc);
        double2 B = make_double2((E.x - O.x) * sc, (E.y - O.y) * sc);
        if (a.job.imul) {
          const double k = a.job.imul == 2 ? (double)m : 1.0;  // i m or i
          A = make_double2(-k * A.y, k * A.x);
          B = make_double2(-k * B.y, k * B.x);
        }
        if (m == 0) {
          A.y = 0.0;
          B.y = 0.0;
        }
        buf[m] = make_double2(A.x - B.y, A.y + B.x);
        if (m > 0) buf[N - m] = make_double2(A.x + B.y, -A.y + B.x);
      }
      wave_sync();
      double2 v[1][L];
      static_for<L>([&](auto n1_) { v[0][n1_] = buf[64 * decltype(n1_)::value + lane]; });
      wave_sync();
      wave_fft64<L, 1, +1>(v, buf, tw, lane);
      float* rn = a.job.grid + (size_t)fld * plane + (size_t)(a.nlat - 1 - i) * N;
      float* rs = a.job.grid + (size_t)fld * plane + (size_t)i * N;
      static_for<L>([&](auto q_) {
        const int k = lane + 64 * decltype(q_)::value;
        const double2 x = buf[k];
        rn[k] = (float)x.x;
        rs[k] = (float)x.y;
      });
    }
    __syncthreads();  // every span is consumed before the next group's runs land
  }
}

// L of the wave-resident path for N, or 0 (generic path).
int wave_L(int N) {
  if (N % 64) return 0;
  const int L = N / 64;
  for (int l : {2, 3, 4, 6, 8, 12, 16, 24})
    if (L == l) return L;
  return 0;
}
// Grid of a persistent wave-resident launch: as many blocks as fit on the
// device at once (registers and LDS, from the runtime's occupancy query, at
// most 8 per CU), capped by the number of groups.  A grid larger than what is
// resident runs its surplus blocks after the first ones finish, which leaves
// the CUs half idle at the end.  The query is cached per (kernel, LDS size).
int persistent_blocks(const void* kernel, int threads, int groups, size_t lds_bytes) {
  static std::mutex mu;
  static std::map<std::pair<const void*, size_t>, int> cache;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    cus = 256;
  int per_cu = 0;
  {
    std::lock_guard<std::mutex> lock(mu);
    const auto key = std::make_pair(kernel, lds_bytes);
    const auto it = cache.find(key);
    if (it != cache.end()) {
      per_cu = it->second;
    } else {
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds_bytes) !=
              hipSuccess || per_cu < 1) {
        (void)hipGetLastError();
        per_cu = std::max(1, (int)((size_t)(160 * 1024) / std::max<size_t>(lds_bytes, 1)));
      }
      per_cu = std::min(per_cu, 8);
      cache.emplace(key, per_cu);
    }
  }
  return std::max(1, std::min(groups, cus * per_cu));
}

// One persistent launch of a wave-resident kernel (rb rows per block).
template <class Args>
hipError_t launch_persistent(void (*kernel)(Args), int groups, int rb, size_t lds, hipStream_t s,
                             const Args& a) {
  const void* k = reinterpret_cast<const void*>(kernel);
  const dim3 g(persistent_blocks(k, 64 * rb, groups, lds)), b(64 * rb);
  Args copy = a;
  void* args[] = {&copy};
  return hipLaunchKernel(k, g, b, args, lds, s);
}

size_t wave64_lds_bytes(int N, int rb) {
  const int L = N / 64;
  return ((size_t)N + 64 + (size_t)rb * (L * kRowSpan + kSpanPad)) * sizeof(double2);
}
size_t fold_wave_lds_bytes(int N, int vector, int rb) {
  const int L = N / 64;
  return ((size_t)N + 64 + (size_t)rb * ((vector ? 2 : 1) * L * kRowSpan + kSpanPad)) * sizeof(float2);
}
size_t unfold_wave_lds_bytes(int N, int rb) {
  const int L = N / 64;
  return ((size_t)N + 64 + (size_t)rb * (L * kRowSpan + kSpanPad)) * sizeof(float2);
}
constexpr size_t kWaveLdsMax = 160 * 1024;  // LDS of one gfx950 workgroup
// rows per block: 8 when their LDS fits, else 4
int wave_rb(size_t lds8) { return lds8 <= kWaveLdsMax ? 8 : 4; }

size_t fold_lds_bytes(int N, int T, int vector, int rp) {
  (void)T;
  return (size_t)(1 + rp * (vector ? 3 : 2)) * N * sizeof(float2);
}
size_t unfold_lds_bytes(int N, int T, int rp) {
  (void)T;
  return (size_t)(1 + rp * 2) * N * sizeof(float2);
}

// ---- the reference's Fourier waist (*lead, nlat, T+1) complex ----
// (FourierWaist, reference transform.py:417-629; fourier_waist :1294-1404,
// _rows_to_grid :1654-1702.)  One wavefront per row, the row alone in the
// FFT (imaginary part zero): a row's spectrum depends on that row only, so a
// band fill or drain writes the bits the whole-field call writes.

constexpr int kRowWaves = 4;  // wavefronts (rows) per block

__device__ __forceinline__ size_t waist_row(const WaistRowsArgs& a, int r) {
  return ((size_t)(r / a.nrows) * a.nlat + a.row0 + r % a.nrows) * (size_t)(a.T + 1);
}

// rows -> waist: X[m] / nlon for m <= T.
template <int L>
__global__ void __launch_bounds__(64 * kRowWaves) waist_from_rows_kernel(WaistRowsArgs a) {
  extern __shared__ float2 lds[];
  constexpr int N = 64 * L, SP = wave_span<L>();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  float2* tw = lds;
  float2* buf = lds + wave_tw<L>() + (size_t)w * SP;
  fill_wave_twiddles<L>(tw, a.fft.tw);
  __syncthreads();
  const int r = blockIdx.x * kRowWaves + w;
  if (r >= a.nlead * a.nrows) return;  // no block barrier below
  const float* x = a.rows_in + (size_t)r * N;
  float2 v[1][L];
  static_for<L>([&](auto n1_) {
    constexpr int n1 = decltype(n1_)::value;
    v[0][n1] = make_float2(x[64 * n1 + lane], 0.f);
  });
  wave_fft64<L, 1, -1>(v, buf, tw, lane);
  const float inv = 1.0f / (float)N;
  float2* out = a.waist_out + waist_row(a, r);
  for (int m = lane; m <= a.T; m += 64) {
    const float2 z = buf[m];
    out[m] = make_float2(z.x * inv, z.y * inv);
  }
}

// waist -> rows: x[k] = Re X[0] + 2 Re sum_{m=1..T} X[m] e^{2 pi i k m / N}
// (irfft of nlon X with the orders above T zero).
template <int L>
__global__ void __launch_bounds__(64 * kRowWaves) rows_from_waist_kernel(WaistRowsArgs a) {
  extern __shared__ float2 lds[];
  constexpr int N = 64 * L, SP = wave_span<L>();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  float2* tw = lds;
  float2* buf = lds + wave_tw<L>() + (size_t)w * SP;
  fill_wave_twiddles<L>(tw, a.fft.tw);
  __syncthreads();
  const int r = blockIdx.x * kRowWaves + w;
  if (r >= a.nlead * a.nrows) return;  // no block barrier below
  const float2* X = a.waist_in + waist_row(a, r);
  for (int k = lane; k < N; k += 64) buf[k] = make_float2(0.f, 0.f);
  wave_sync();
  for (int m = lane; m <= a.T; m += 64) {
    const float2 A = X[m];
    if (m == 0) {
      buf[0] = make_float2(A.x, 0.f);  // irfft reads the real part of the m = 0 bin
    } else {
      buf[m] = A;
      buf[N - m] = cconj(A);
    }
  }
  wave_sync();
  float2 v[1][L];
  static_for<L>([&](auto n1_) { v[0][n1_] = buf[64 * decltype(n1_)::value + lane]; });
  wave_sync();
  wave_fft64<L, 1, +1>(v, buf, tw, lane);
  float* out = a.rows_out + (size_t)r * N;
  for (int k = lane; k < N; k += 64) out[k] = buf[k].x;
}

// The same two maps for an nlon that is not 64 L: direct sums against the
// twiddle table, one thread per output value.
__global__ void waist_from_rows_dft_kernel(WaistRowsArgs a) {
  const int N = a.fft.N, T1 = a.T + 1;
  const long long total = (long long)a.nlead * a.nrows * T1;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(e / T1), m = (int)(e % T1);
    const float* x = a.rows_in + (size_t)r * N;
    float2 acc = make_float2(0.f, 0.f);
    for (int j = 0; j < N; ++j) {
      const float2 t = a.fft.tw[(int)(((long long)j * m) % N)];  // e^{-2 pi i j m / N}
      acc.x += x[j] * t.x;
      acc.y += x[j] * t.y;
    }
    a.waist_out[waist_row(a, r) + m] = make_float2(acc.x / (float)N, acc.y / (float)N);
  }
}

__global__ void rows_from_waist_dft_kernel(WaistRowsArgs a) {
  const int N = a.fft.N, T1 = a.T + 1;
  const long long total = (long long)a.nlead * a.nrows * N;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(e / N), k = (int)(e % N);
    const float2* X = a.waist_in + waist_row(a, r);
    float acc = X[0].x;
    for (int m = 1; m < T1; ++m) {
      const float2 t = a.fft.tw[(int)(((long long)k * m) % N)];  // conj: e^{+2 pi i k m / N}
      acc += 2.f * (X[m].x * t.x + X[m].y * t.y);
    }
    a.rows_out[(size_t)r * N + k] = acc;
  }
}

// reference waist -> folded analysis waist: one thread per (field, folded row
// i < JHP, order m < 4 ceil((T+1)/4)); padding rows and orders are zeros.
__global__ void waist_fold_kernel(WaistFoldArgs a) {
  const int T1 = a.T + 1, MO = (T1 + kMT - 1) / kMT * kMT;
  const long long total = (long long)a.nfields * a.JHP * MO;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int m = (int)(e % MO);
    const int i = (int)((e / MO) % a.JHP);
    const int f = (int)(e / ((long long)MO * a.JHP));
    const bool live = i < a.JH && m < T1;
    const float s = live ? a.rowscale[i] : 0.f;
    const size_t rn = ((size_t)f * a.nlat + (a.nlat - 1 - (live ? i : 0))) * T1 + (live ? m : 0);
    const size_t rs = ((size_t)f * a.nlat + (live ? i : 0)) * T1 + (live ? m : 0);
    float2 e0[2], o0[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      e0[k] = make_float2(0.f, 0.f);
      o0[k] = make_float2(0.f, 0.f);
      const float2* W = k == 0 ? a.w0 : a.w1;
      if (live && (k == 0 || a.vector)) {
        const float2 xn = W[rn], xs = W[rs];
        e0[k] = make_float2((xn.x + xs.x) * s, (xn.y + xs.y) * s);
        o0[k] = make_float2((xn.x - xs.x) * s, (xn.y - xs.y) * s);
      }
    }
    if (!a.vector) {
      const int c = a.col0 + 2 * f;
      a.waist[waist_index(c, a.ldw, a.JHP, m, 0, i)] = e0[0].x;
      a.waist[waist_index(c, a.ldw, a.JHP, m, 1, i)] = o0[0].x;
      a.waist[waist_index(c + 1, a.ldw, a.JHP, m, 0, i)] = e0[0].y;
      a.waist[waist_index(c + 1, a.ldw, a.JHP, m, 1, i)] = o0[0].y;
    } else if (a.vector == 2) {
      // divergence alone: iU = [-U.im, U.re] against P m / cos, -V against dP/dphi
      const int c = a.col0 + 2 * f;
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const float2 U = p == 0 ? e0[0] : o0[0];
        const float2 V = p == 0 ? e0[1] : o0[1];
        a.waist[waist_index(c, a.ldw, a.JHP, m, p, i)] = -U.y;
        a.waist[waist_index(c + 1, a.ldw, a.JHP, m, p, i)] = U.x;
        a.waist1[waist_index(c, a.ldw, a.JHP, m, p, i)] = -V.x;
        a.waist1[waist_index(c + 1, a.ldw, a.JHP, m, p, i)] = -V.y;
      }
    } else {
      // level columns [iV, iU] = [-V.im, V.re, -U.im, U.re] (w0 = u, w1 = v)
      const int c = a.col0 + 4 * f;
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const float2 U = p == 0 ? e0[0] : o0[0];
        const float2 V = p == 0 ? e0[1] : o0[1];
        a.waist[waist_index(c + 0, a.ldw, a.JHP, m, p, i)] = -V.y;
        a.waist[waist_index(c + 1, a.ldw, a.JHP, m, p, i)] = V.x;
        a.waist[waist_index(c + 2, a.ldw, a.JHP, m, p, i)] = -U.y;
        a.waist[waist_index(c + 3, a.ldw, a.JHP, m, p, i)] = U.x;
      }
    }
  }
}

// folded synthesis waist -> reference waist; one thread per (field, folded
// row i < JH, order m <= T).
__global__ void waist_unfold_kernel(WaistUnfoldArgs a) {
  const int T1 = a.T + 1;
  const long long total = (long long)a.nfields * a.JH * T1;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int m = (int)(e % T1);
    const int i = (int)((e / T1) % a.JH);
    const int f = (int)(e / ((long long)T1 * a.JH));
    const int c = a.col0 + a.fstride * f;
    const float2 E = make_float2(a.waist[waist_index(c, a.ldw, a.JHP, m, 0, i)],
                                 a.waist[waist_index(c + 1, a.ldw, a.JHP, m, 0, i)]);
    const float2 O = make_float2(a.waist[waist_index(c, a.ldw, a.JHP, m, 1, i)],
                                 a.waist[waist_index(c + 1, a.ldw, a.JHP, m, 1, i)]);
    const int north = a.nlat - 1 - i;
    a.out[((size_t)f * a.nlat + north) * T1 + m] = make_float2(E.x + O.x, E.y + O.y);
    if (north != i) a.out[((size_t)f * a.nlat + i) * T1 + m] = make_float2(E.x - O.x, E.y - O.y);
  }
}

size_t rows_wave_lds_bytes(int N) {
  const int L = N / 64;
  return ((size_t)N + 64 + (size_t)kRowWaves * L * kRowSpan) * sizeof(float2);
}

// Register prefetch depth per lane: PF * 64 >= N, or 0 (no prefetch) above 1536.
int fft_pf(int N) {
  const int need = (N + kWave - 1) / kWave;
  const int opts[] = {2, 4, 8, 12, 16, 24};
  for (int o : opts)
    if (need <= o) return o;
  return 0;
}

}  // namespace sht

#define SHT_PF_DISPATCH(KERNEL, PF, ...)                                 \
  switch (PF) {                                                          \
    case 2: hipLaunchKernelGGL(sht::KERNEL<2>, __VA_ARGS__); break;      \
    case 4: hipLaunchKernelGGL(sht::KERNEL<4>, __VA_ARGS__); break;      \
    case 8: hipLaunchKernelGGL(sht::KERNEL<8>, __VA_ARGS__); break;      \
    case 12: hipLaunchKernelGGL(sht::KERNEL<12>, __VA_ARGS__); break;    \
    case 16: hipLaunchKernelGGL(sht::KERNEL<16>, __VA_ARGS__); break;    \
    case 24: hipLaunchKernelGGL(sht::KERNEL<24>, __VA_ARGS__); break;    \
    default: hipLaunchKernelGGL(sht::KERNEL<0>, __VA_ARGS__); break;     \
  }

#define SHT_L_DISPATCH(L, LAUNCH) \
  switch (L) {                     \
    case 2: LAUNCH(2); break;      \
    case 3: LAUNCH(3); break;      \
    case 4: LAUNCH(4); break;      \
    case 6: LAUNCH(6); break;      \
    case 8: LAUNCH(8); break;      \
    case 12: LAUNCH(12); break;    \
    case 16: LAUNCH(16); break;    \
    case 24: LAUNCH(24); break;    \
    default: return hipErrorInvalidValue; \
  }

extern "C" hipError_t sht_launch_rfft_fold(const sht::FoldArgs* a, hipStream_t s) {
  if (a->job.low) {  // FP64 fold of one field of a vector pair (wave-resident N only)
    const int L = a->fft.wave_L;
    if (!L || !a->fft.tw64) return hipErrorInvalidValue;
    const int rb = sht::wave_rb(sht::wave64_lds_bytes(a->fft.N, 8));
    const int groups = a->job.nfields * (a->JHP / rb);
    if (groups == 0) return hipSuccess;
    const size_t lds = sht::wave64_lds_bytes(a->fft.N, rb);
    hipError_t e = hipSuccess;
#define SHT_FOLD64(LL)                                                                   \
  e = sht::launch_persistent(rb == 8 ? sht::rfft_fold_wave64_kernel<LL, 8>               \
                                     : sht::rfft_fold_wave64_kernel<LL, 4>,              \
                             groups, rb, lds, s, *a)
    SHT_L_DISPATCH(L, SHT_FOLD64)
#undef SHT_FOLD64
    return e;
  }
  if (const int L = a->fft.wave_L) {
    const int rb = sht::wave_rb(sht::fold_wave_lds_bytes(a->fft.N, a->job.vector, 8));
    const int groups = a->job.nfields * (a->JHP / rb);
    if (groups == 0) return hipSuccess;
    const size_t lds = sht::fold_wave_lds_bytes(a->fft.N, a->job.vector, rb);
    hipError_t e = hipSuccess;
#define SHT_FOLD(LL)                                                                          \
  e = sht::launch_persistent(a->job.vector ? (rb == 8 ? sht::rfft_fold_wave_kernel<LL, 2, 8>   \
                                                      : sht::rfft_fold_wave_kernel<LL, 2, 4>)  \
                                           : (rb == 8 ? sht::rfft_fold_wave_kernel<LL, 1, 8>   \
                                                      : sht::rfft_fold_wave_kernel<LL, 1, 4>), \
                             groups, rb, lds, s, *a)
    SHT_L_DISPATCH(L, SHT_FOLD)
#undef SHT_FOLD
    return e;
  }
  size_t lds = sht::fold_lds_bytes(a->fft.N, a->T, a->job.vector, a->rp);
  int blocks = a->nchunks * a->job.nfields;
  if (blocks == 0) return hipSuccess;
  if (a->job.rowscale_hi) return hipErrorInvalidValue;  // order-dependent scales: wave path only
  SHT_PF_DISPATCH(rfft_fold_kernel, sht::fft_pf(a->fft.N), dim3(blocks), dim3(sht::kWave * a->rp),
                  lds, s, *a);
  return hipGetLastError();
}

extern "C" hipError_t sht_launch_irfft_unfold(const sht::UnfoldArgs* a, hipStream_t s) {
  if (a->job.f64) {  // FP64 unfold (wave-resident N only)
    const int L = a->fft.wave_L;
    if (!L || !a->fft.tw64 || !a->job.low || !a->job.rowscale64) return hipErrorInvalidValue;
    const int rb = sht::wave_rb(sht::wave64_lds_bytes(a->fft.N, 8));
    const int groups = a->job.nfields * ((a->JH + rb - 1) / rb);
    if (groups == 0) return hipSuccess;
    const size_t lds = sht::wave64_lds_bytes(a->fft.N, rb);
    hipError_t e = hipSuccess;
#define SHT_UNFOLD64(LL)                                                                 \
  e = sht::launch_persistent(rb == 8 ? sht::irfft_unfold_wave64_kernel<LL, 8>            \
                                     : sht::irfft_unfold_wave64_kernel<LL, 4>,           \
                             groups, rb, lds, s, *a)
    SHT_L_DISPATCH(L, SHT_UNFOLD64)
#undef SHT_UNFOLD64
    return e;
  }
  if (const int L = a->fft.wave_L) {
    const int rb = sht::wave_rb(sht::unfold_wave_lds_bytes(a->fft.N, 8));
    const int groups = a->job.nfields * ((a->JH + rb - 1) / rb);
    if (groups == 0) return hipSuccess;
    const size_t lds = sht::unfold_wave_lds_bytes(a->fft.N, rb);
    hipError_t e = hipSuccess;
#define SHT_UNFOLD(LL)                                                                       \
  e = sht::launch_persistent(rb == 8 ? sht::irfft_unfold_wave_kernel<LL, 8>                 \
                                     : sht::irfft_unfold_wave_kernel<LL, 4>,                \
                             groups, rb, lds, s, *a)
    SHT_L_DISPATCH(L, SHT_UNFOLD)
#undef SHT_UNFOLD
    return e;
  }
  size_t lds = sht::unfold_lds_bytes(a->fft.N, a->T, a->rp);
  int blocks = a->nchunks * a->job.nfields;
  if (blocks == 0) return hipSuccess;
  if (a->job.rowscale_hi) return hipErrorInvalidValue;  // order-dependent scales: wave path only
  SHT_PF_DISPATCH(irfft_unfold_kernel, sht::fft_pf(a->fft.N), dim3(blocks),
                  dim3(sht::kWave * a->rp), lds, s, *a);
  return hipGetLastError();
}

template <int PF>
static hipError_t allow_lds_one(size_t bytes) {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(sht::rfft_fold_kernel<PF>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(sht::irfft_unfold_kernel<PF>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <int L>
static hipError_t allow_lds_wave(size_t bytes) {
  for (const void* k : {reinterpret_cast<const void*>(sht::rfft_fold_wave_kernel<L, 1, 4>),
                        reinterpret_cast<const void*>(sht::rfft_fold_wave_kernel<L, 2, 4>),
                        reinterpret_cast<const void*>(sht::rfft_fold_wave_kernel<L, 1, 8>),
                        reinterpret_cast<const void*>(sht::rfft_fold_wave_kernel<L, 2, 8>),
                        reinterpret_cast<const void*>(sht::irfft_unfold_wave_kernel<L, 4>),
                        reinterpret_cast<const void*>(sht::irfft_unfold_wave_kernel<L, 8>),
                        reinterpret_cast<const void*>(sht::waist_from_rows_kernel<L>),
                        reinterpret_cast<const void*>(sht::rows_from_waist_kernel<L>),
                        reinterpret_cast<const void*>(sht::rfft_fold_wave64_kernel<L, 4>),
                        reinterpret_cast<const void*>(sht::rfft_fold_wave64_kernel<L, 8>),
                        reinterpret_cast<const void*>(sht::irfft_unfold_wave64_kernel<L, 4>),
                        reinterpret_cast<const void*>(sht::irfft_unfold_wave64_kernel<L, 8>)}) {
    hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

extern "C" int sht_fft_wave_L(int N) { return sht::wave_L(N); }
// LDS of the FP64 fold and unfold launches (8 rows per block where they fit, else 4)
extern "C" size_t sht_fold64_lds_bytes(int N) {
  const size_t b8 = sht::wave64_lds_bytes(N, 8);
  return sht::wave_rb(b8) == 8 ? b8 : sht::wave64_lds_bytes(N, 4);
}
extern "C" size_t sht_unfold64_lds_bytes(int N) { return sht_fold64_lds_bytes(N); }
extern "C" size_t sht_fft_wave_lds_bytes(int N, int vector) {
  // the LDS a launch asks for: 8 rows per block where that fits, else 4
  if (vector >= 0) {
    const size_t b8 = sht::fold_wave_lds_bytes(N, vector, 8);
    return sht::wave_rb(b8) == 8 ? b8 : sht::fold_wave_lds_bytes(N, vector, 4);
  }
  const size_t b8 = sht::unfold_wave_lds_bytes(N, 8);
  return sht::wave_rb(b8) == 8 ? b8 : sht::unfold_wave_lds_bytes(N, 4);
}

extern "C" hipError_t sht_fft_allow_lds(size_t bytes, size_t wave_bytes) {
  hipError_t e = hipSuccess;
  const size_t wb = wave_bytes;
  for (hipError_t r : {allow_lds_wave<2>(wb), allow_lds_wave<3>(wb), allow_lds_wave<4>(wb),
                       allow_lds_wave<6>(wb), allow_lds_wave<8>(wb), allow_lds_wave<12>(wb),
                       allow_lds_wave<16>(wb), allow_lds_wave<24>(wb)})
    if (r != hipSuccess) e = r;
  for (hipError_t r : {allow_lds_one<0>(bytes), allow_lds_one<2>(bytes), allow_lds_one<4>(bytes),
                       allow_lds_one<8>(bytes), allow_lds_one<12>(bytes), allow_lds_one<16>(bytes),
                       allow_lds_one<24>(bytes)})
    if (r != hipSuccess) e = r;
  return e;
}

extern "C" size_t sht_fold_lds_bytes(int N, int T, int vector, int rp) {
  return sht::fold_lds_bytes(N, T, vector, rp);
}
extern "C" size_t sht_unfold_lds_bytes(int N, int T, int rp) { return sht::unfold_lds_bytes(N, T, rp); }

static unsigned grid_for(long long total) {
  long long b = (total + 255) / 256;
  if (b > 65536) b = 65536;
  return (unsigned)(b < 1 ? 1 : b);
}

extern "C" hipError_t sht_launch_waist_from_rows(const sht::WaistRowsArgs* a, hipStream_t s) {
  const long long R = (long long)a->nlead * a->nrows;
  if (R <= 0) return hipSuccess;
  if (const int L = a->fft.wave_L) {
    const dim3 g((unsigned)((R + sht::kRowWaves - 1) / sht::kRowWaves)), b(64 * sht::kRowWaves);
    const size_t lds = sht::rows_wave_lds_bytes(a->fft.N);
#define SHT_W2R(LL) hipLaunchKernelGGL((sht::waist_from_rows_kernel<LL>), g, b, lds, s, *a)
    SHT_L_DISPATCH(L, SHT_W2R)
#undef SHT_W2R
  } else {
    hipLaunchKernelGGL(sht::waist_from_rows_dft_kernel, dim3(grid_for(R * (a->T + 1))), dim3(256),
                       0, s, *a);
  }
  return hipGetLastError();
}

extern "C" hipError_t sht_launch_rows_from_waist(const sht::WaistRowsArgs* a, hipStream_t s) {
  const long long R = (long long)a->nlead * a->nrows;
  if (R <= 0) return hipSuccess;
  if (const int L = a->fft.wave_L) {
    const dim3 g((unsigned)((R + sht::kRowWaves - 1) / sht::kRowWaves)), b(64 * sht::kRowWaves);
    const size_t lds = sht::rows_wave_lds_bytes(a->fft.N);
#define SHT_R2W(LL) hipLaunchKernelGGL((sht::rows_from_waist_kernel<LL>), g, b, lds, s, *a)
    SHT_L_DISPATCH(L, SHT_R2W)
#undef SHT_R2W
  } else {
    hipLaunchKernelGGL(sht::rows_from_waist_dft_kernel, dim3(grid_for(R * a->fft.N)), dim3(256),
                       0, s, *a);
  }
  return hipGetLastError();
}

extern "C" hipError_t sht_launch_waist_fold(const sht::WaistFoldArgs* a, hipStream_t s) {
  const int MO = (a->T + 1 + sht::kMT - 1) / sht::kMT * sht::kMT;
  const long long total = (long long)a->nfields * a->JHP * MO;
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(sht::waist_fold_kernel, dim3(grid_for(total)), dim3(256), 0, s, *a);
  return hipGetLastError();
}

extern "C" hipError_t sht_launch_waist_unfold(const sht::WaistUnfoldArgs* a, hipStream_t s) {
  const long long total = (long long)a->nfields * a->JH * (a->T + 1);
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(sht::waist_unfold_kernel, dim3(grid_for(total)), dim3(256), 0, s, *a);
  return hipGetLastError();
}
