// Longitude real FFTs fused with the parity fold (analysis) and unfold
// (synthesis).
//
// The waist is stored m-tiled:  W[c][m/8][p][i][m%8]  (fp32), per data
// column c (2*field + re/im).  A block holds `rp` wavefronts; wavefront w owns
// folded row i0 + w (the mirror pair north = nlat-1-i, south = i) and runs
// its own FFT in its own LDS slice, so rp FFTs advance together between the
// block barriers of the Stockham passes.  Neighbouring wavefronts write
// neighbouring 32-byte runs of each m-tile, which meet in L2 as 128-byte
// lines.
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
#include "sht_kernels.h"

namespace sht {

constexpr int kWave = 64;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// Twiddle e^{sign * 2 pi i k / N} from the forward table.
__device__ __forceinline__ float2 twiddle(const float2* tw, int k, int sign) {
  float2 t = tw[k];
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
  auto put = [&](int g, int src, float2* b) {
    const int i = g * a.rp + w;
    if (i >= a.JH) return;
    if constexpr (PF > 0) {
#pragma unroll
      for (int j = 0; j < PF; ++j) {
        const int k = lane + kWave * j;
        if (k < N) b[k] = make_float2(rn[src][j], rs[src][j]);
      }
    } else {
      const float* n_ = gsrc[src] + (size_t)(a.nlat - 1 - i) * N;
      const float* s_ = gsrc[src] + (size_t)i * N;
      for (int k = lane; k < N; k += kWave) b[k] = make_float2(n_[k], s_[k]);
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
      if (!a.job.vector) {
        const int c = a.job.col0 + 2 * item;
        a.job.waist0[waist_index(c, a.job.ldw, a.JHP, m, 0, i)] = e[0].x;
        a.job.waist0[waist_index(c + 1, a.job.ldw, a.JHP, m, 0, i)] = e[0].y;
        a.job.waist0[waist_index(c, a.job.ldw, a.JHP, m, 1, i)] = o[0].x;
        a.job.waist0[waist_index(c + 1, a.job.ldw, a.JHP, m, 1, i)] = o[0].y;
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
          for (int cl = 0; cl < 4; ++cl)
            a.job.waist0[waist_index(c + cl, a.job.ldw, a.JHP, m, p, i)] = va[cl];
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
      A = make_float2(-A.y, A.x);
      B = make_float2(-B.y, B.x);
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
        rn[k] = v.x;
        rs[k] = v.y;
      }
    }
    __syncthreads();  // buffers are rewritten by the next group
  }
}

size_t fold_lds_bytes(int N, int T, int vector, int rp) {
  (void)T;
  return (size_t)(1 + rp * (vector ? 3 : 2)) * N * sizeof(float2);
}
size_t unfold_lds_bytes(int N, int T, int rp) {
  (void)T;
  return (size_t)(1 + rp * 2) * N * sizeof(float2);
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

extern "C" hipError_t sht_launch_rfft_fold(const sht::FoldArgs* a, hipStream_t s) {
  size_t lds = sht::fold_lds_bytes(a->fft.N, a->T, a->job.vector, a->rp);
  int blocks = a->nchunks * a->job.nfields;
  if (blocks == 0) return hipSuccess;
  SHT_PF_DISPATCH(rfft_fold_kernel, sht::fft_pf(a->fft.N), dim3(blocks), dim3(sht::kWave * a->rp),
                  lds, s, *a);
  return hipGetLastError();
}

extern "C" hipError_t sht_launch_irfft_unfold(const sht::UnfoldArgs* a, hipStream_t s) {
  size_t lds = sht::unfold_lds_bytes(a->fft.N, a->T, a->rp);
  int blocks = a->nchunks * a->job.nfields;
  if (blocks == 0) return hipSuccess;
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

extern "C" hipError_t sht_fft_allow_lds(size_t bytes) {
  hipError_t e = hipSuccess;
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
