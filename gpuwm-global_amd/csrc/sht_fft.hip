// Longitude real FFTs fused with the parity fold (analysis) and unfold
// (synthesis).
//
// The waist is stored m-tiled:  W[c][m/8][p][i][m%8]  (fp32), per data
// column c (2*field + re/im), so that four consecutive folded rows of one
// m-tile are one 128-byte run.  A block handles rp (<= 4) consecutive folded
// rows of one field (or of one u/v pair): it stages their waist tile in LDS
// and runs one complex FFT per mirror pair.
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


__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// Twiddle e^{sign * 2 pi i k / N} from the forward table.
__device__ __forceinline__ float2 twiddle(const float2* tw, int k, int N, int sign) {
  float2 t = tw[k % N];
  return sign < 0 ? t : cconj(t);
}

// One Stockham pass of radix R (compile-time for 2/3/4, runtime otherwise).
template <int R>
__device__ __forceinline__ void dft_small(float2* v, int sign) {
  if constexpr (R == 2) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    float2 a2 = cadd(v[1], v[3]), a3 = csub(v[1], v[3]);
    // multiply a3 by -i (forward) or +i (inverse)
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

template <int R>
__device__ void stockham_pass(const float2* __restrict__ in, float2* __restrict__ out, int N,
                              int Ns, const float2* tw, int sign) {
  const int nb = N / R;
  for (int j = threadIdx.x; j < nb; j += blockDim.x) {
    const int k = j % Ns;
    const int tstep = N / (Ns * R);
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      float2 x = in[j + r * nb];
      v[r] = (r == 0 || k == 0) ? x : cmul(x, twiddle(tw, r * k * tstep, N, sign));
    }
    dft_small<R>(v, sign);
    const int d = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) out[d + r * Ns] = v[r];
  }
}

__device__ void stockham_pass_generic(const float2* __restrict__ in, float2* __restrict__ out,
                                      int N, int Ns, int R, const float2* tw, int sign) {
  const int nb = N / R;
  const int rstep = N / R;  // twiddle stride of the R-point DFT
  for (int j = threadIdx.x; j < nb; j += blockDim.x) {
    const int k = j % Ns;
    const int tstep = N / (Ns * R);
    const int d = (j / Ns) * Ns * R + k;
    for (int q = 0; q < R; ++q) {
      float2 acc = make_float2(0.f, 0.f);
      for (int r = 0; r < R; ++r) {
        float2 x = in[j + r * nb];
        if (r > 0 && k > 0) x = cmul(x, twiddle(tw, r * k * tstep, N, sign));
        if (r > 0 && q > 0) x = cmul(x, twiddle(tw, ((r * q) % R) * rstep, N, sign));
        acc = cadd(acc, x);
      }
      out[d + q * Ns] = acc;
    }
  }
}

// Full FFT of N points held in buf[0]; returns the buffer holding the result.
__device__ float2* block_fft(float2* buf0, float2* buf1, const FftPlan& plan, int sign) {
  float2* in = buf0;
  float2* out = buf1;
  int Ns = 1;
  for (int s = 0; s < plan.nstages; ++s) {
    const int R = plan.radix[s];
    if (R == 4) stockham_pass<4>(in, out, plan.N, Ns, plan.tw, sign);
    else if (R == 2) stockham_pass<2>(in, out, plan.N, Ns, plan.tw, sign);
    else if (R == 3) stockham_pass<3>(in, out, plan.N, Ns, plan.tw, sign);
    else stockham_pass_generic(in, out, plan.N, Ns, R, plan.tw, sign);
    __syncthreads();
    Ns *= R;
    float2* t = in;
    in = out;
    out = t;
  }
  return in;
}

// ---------------------------------------------------------------------------
// Analysis: grid rows -> folded, weighted waist.
// ---------------------------------------------------------------------------




// Fold-stage LDS: two FFT buffers of N points and, per row of the group,
// E/O for m = 0..T: [field-in-pair][p][q][m] complex.
__global__ void __launch_bounds__(kFftThreads)
rfft_fold_kernel(FoldArgs a) {
  extern __shared__ float2 lds[];
  const int N = a.fft.N;
  const int T1 = a.T + 1;
  const int group = blockIdx.x % a.ngroups;
  const int item = blockIdx.x / a.ngroups;  // field (scalar) or level (vector)
  if (item >= a.job.nfields) return;
  const int nsrc = a.job.vector ? 2 : 1;
  float2* buf0 = lds;
  float2* buf1 = lds + N;
  float2* eo = lds + 2 * N;  // [src][p][q][m]
  const int i0 = group * a.rp;
  const size_t plane = (size_t)a.nlat * N;

  for (int src = 0; src < nsrc; ++src) {
    const float* g = (src == 0 ? a.job.grid0 : a.job.grid1) + (size_t)item * plane;
    for (int q = 0; q < a.rp; ++q) {
      const int i = i0 + q;
      float2* eoq = eo + ((size_t)(src * 2) * a.rp + q) * T1;
      float2* eoq1 = eo + ((size_t)(src * 2 + 1) * a.rp + q) * T1;
      if (i >= a.JH) {  // padding rows of the waist are written as zeros
        for (int m = threadIdx.x; m < T1; m += blockDim.x) {
          eoq[m] = make_float2(0.f, 0.f);
          eoq1[m] = make_float2(0.f, 0.f);
        }
        continue;
      }
      const float* rn = g + (size_t)(a.nlat - 1 - i) * N;
      const float* rs = g + (size_t)i * N;
      for (int k = threadIdx.x; k < N; k += blockDim.x) buf0[k] = make_float2(rn[k], rs[k]);
      __syncthreads();
      float2* Z = block_fft(buf0, buf1, a.fft, -1);
      const float s = a.job.rowscale[i];
      for (int m = threadIdx.x; m < T1; m += blockDim.x) {
        float2 zm = Z[m];
        float2 zc = cconj(Z[(N - m) % N]);
        // X_N = (zm + zc)/2, X_S = (zm - zc)/(2i)
        float2 xn = make_float2(0.5f * (zm.x + zc.x), 0.5f * (zm.y + zc.y));
        float2 d = csub(zm, zc);
        float2 xs = make_float2(0.5f * d.y, -0.5f * d.x);
        eoq[m] = make_float2((xn.x + xs.x) * s, (xn.y + xs.y) * s);
        eoq1[m] = make_float2((xn.x - xs.x) * s, (xn.y - xs.y) * s);
      }
      __syncthreads();
    }
  }
  // Write the group's tile: W[c][mb][p][i0..i0+3][0..7].
  const int MB = (T1 + kMT - 1) / kMT;
  const int ncolv = a.job.vector ? 4 : 2;
  const int per_c = MB * 2 * a.rp * kMT;  // floats per column in this tile
  const int narr = a.job.vector ? 2 : 1;
  for (int arr = 0; arr < narr; ++arr) {
    float* W = arr == 0 ? a.job.waist0 : a.job.waist1;
    for (int e = threadIdx.x; e < ncolv * per_c; e += blockDim.x) {
      const int cl = e / per_c;
      int r = e % per_c;
      const int mm = r % kMT;
      r /= kMT;
      const int q = r % a.rp;
      r /= a.rp;
      const int p = r % 2;
      const int mb = r / 2;
      const int m = mb * kMT + mm;
      float val = 0.f;
      if (m < T1) {
        if (!a.job.vector) {
          float2 v = eo[((size_t)p * a.rp + q) * T1 + m];
          val = (cl == 0) ? v.x : v.y;
        } else {
          // data columns of level l: arrangement a = [-V.im, V.re, -U.im, U.re] (fold p),
          // arrangement b = [U.re, U.im, -V.re, -V.im] (fold p).
          float2 U = eo[((size_t)(0 * 2 + p) * a.rp + q) * T1 + m];
          float2 V = eo[((size_t)(1 * 2 + p) * a.rp + q) * T1 + m];
          if (arr == 0) {
            val = cl == 0 ? -V.y : cl == 1 ? V.x : cl == 2 ? -U.y : U.x;
          } else {
            val = cl == 0 ? U.x : cl == 1 ? U.y : cl == 2 ? -V.x : -V.y;
          }
        }
      }
      const size_t c = (size_t)a.job.col0 + (size_t)item * ncolv + cl;
      W[c * a.job.ldw + ((size_t)(mb * 2 + p) * a.JHP + i0 + q) * kMT + mm] = val;
    }
  }
}

// ---------------------------------------------------------------------------
// Synthesis: folded waist -> grid rows.
// ---------------------------------------------------------------------------




__global__ void __launch_bounds__(kFftThreads)
irfft_unfold_kernel(UnfoldArgs a) {
  extern __shared__ float2 lds[];
  const int N = a.fft.N;
  const int T1 = a.T + 1;
  const int group = blockIdx.x % a.ngroups;
  const int f = blockIdx.x / a.ngroups;
  if (f >= a.job.nfields) return;
  float2* buf0 = lds;
  float2* buf1 = lds + N;
  float2* eo = lds + 2 * N;  // [p][q][m]
  const int i0 = group * a.rp;
  const int MB = (T1 + kMT - 1) / kMT;
  // Stage the waist tile (coalesced 128-byte runs).
  const int per_c = MB * 2 * a.rp * kMT;
  for (int e = threadIdx.x; e < 2 * per_c; e += blockDim.x) {
    const int cl = e / per_c;
    int r = e % per_c;
    const int mm = r % kMT;
    r /= kMT;
    const int q = r % a.rp;
    r /= a.rp;
    const int p = r % 2;
    const int mb = r / 2;
    const int m = mb * kMT + mm;
    const size_t c = (size_t)a.job.col0 + (size_t)a.job.fstride * f + cl;
    float v = a.job.waist[c * a.job.ldw + ((size_t)(mb * 2 + p) * a.JHP + i0 + q) * kMT + mm];
    if (m < T1) {
      float* dst = reinterpret_cast<float*>(&eo[((size_t)p * a.rp + q) * T1 + m]);
      dst[cl] = v;
    }
  }
  __syncthreads();
  const size_t plane = (size_t)a.nlat * N;
  for (int q = 0; q < a.rp; ++q) {
    const int i = i0 + q;
    if (i >= a.JH) break;
    const float s = a.job.rowscale[i];
    // A = north spectrum (E + O), B = south spectrum (E - O), m <= T, zero above.
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
      float2 z = make_float2(0.f, 0.f);
      int m = k;
      bool lower = true;
      if (k > N / 2) { m = N - k; lower = false; }
      if (m < T1) {
        float2 E = eo[((size_t)0 * a.rp + q) * T1 + m];
        float2 O = eo[((size_t)1 * a.rp + q) * T1 + m];
        float2 A = make_float2((E.x + O.x) * s, (E.y + O.y) * s);
        float2 B = make_float2((E.x - O.x) * s, (E.y - O.y) * s);
        if (a.job.imul) {
          A = make_float2(-A.y, A.x);
          B = make_float2(-B.y, B.x);
        }
        if (m == 0) { A.y = 0.f; B.y = 0.f; }
        if (lower) {
          z = make_float2(A.x - B.y, A.y + B.x);  // A + iB
        } else {
          z = make_float2(A.x + B.y, -A.y + B.x); // conj(A) + i conj(B)
        }
      }
      buf0[k] = z;
    }
    __syncthreads();
    float2* x = block_fft(buf0, buf1, a.fft, +1);
    float* rn = a.job.grid + (size_t)f * plane + (size_t)(a.nlat - 1 - i) * N;
    float* rs = a.job.grid + (size_t)f * plane + (size_t)i * N;
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
      float2 v = x[k];
      rn[k] = v.x;
      rs[k] = v.y;
    }
    __syncthreads();
  }
}

size_t fold_lds_bytes(int N, int T, int vector, int rp) {
  return (size_t)(2 * N + (vector ? 2 : 1) * 2 * rp * (T + 1)) * sizeof(float2);
}
size_t unfold_lds_bytes(int N, int T, int rp) {
  return (size_t)(2 * N + 2 * rp * (T + 1)) * sizeof(float2);
}

}  // namespace sht

extern "C" hipError_t sht_launch_rfft_fold(const sht::FoldArgs* a, hipStream_t s) {
  size_t lds = sht::fold_lds_bytes(a->fft.N, a->T, a->job.vector, a->rp);
  int blocks = a->ngroups * a->job.nfields;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(sht::rfft_fold_kernel, dim3(blocks), dim3(sht::kFftThreads), lds, s, *a);
  return hipGetLastError();
}

extern "C" hipError_t sht_launch_irfft_unfold(const sht::UnfoldArgs* a, hipStream_t s) {
  size_t lds = sht::unfold_lds_bytes(a->fft.N, a->T, a->rp);
  int blocks = a->ngroups * a->job.nfields;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(sht::irfft_unfold_kernel, dim3(blocks), dim3(sht::kFftThreads), lds, s, *a);
  return hipGetLastError();
}

extern "C" size_t sht_fold_lds_bytes(int N, int T, int vector, int rp) {
  return sht::fold_lds_bytes(N, T, vector, rp);
}
extern "C" size_t sht_unfold_lds_bytes(int N, int T, int rp) { return sht::unfold_lds_bytes(N, T, rp); }
