// Shared definitions for the MI355X (gfx950) spherical-harmonic transform engine.
//
// Layout vocabulary (all device buffers, all little-endian):
//   T        truncation; orders m = 0..T, degrees n = m..T.
//   nlat     Gaussian latitudes, ascending (row 0 is the southernmost).
//   JH       folded latitude rows = ceil(nlat/2); folded row i pairs the
//            north row nlat-1-i with the south row i (odd nlat: the
//            equator row is its own mirror, i = JH-1).
//   parity   p = (n - m) mod 2.  Pbar_n^m(-x) = (-1)^p Pbar_n^m(x), so a
//            table needs only the north nodes and the even/odd sums.
//   K_p(m)   degrees of order m with parity p: ceil((T+1-m-p)/2).
//   KP(m,p)  K_p(m) rounded up to a multiple of 32 (zero padded).
//   JHP      JH rounded up to a multiple of 32 (zero padded).
//
// Legendre table block (m, p), "synthesis" layout: A[r = i][k = n_idx],
// JHP rows x KP(m,p) columns, k contiguous.  "Analysis" layout:
// A[r = n_idx][k = i], KP(m,p) rows x JHP columns, k contiguous.  Both
// layouts hold the same numbers; blocks are stored back to back in
// (m, p) order.  n = m + p + 2*n_idx.
//
// Folded waist (one per data column c = 2*field + {0 re, 1 im}):
//   W[c][m][p][i]  -> offset c*ldw + (2m+p)*JHP + i,  ldw = 2(T+1)*JHP.
// Packed coefficients (one per data column):
//   C[c][m][p][n_idx] -> offset c*ldc + koff(m,p) + n_idx,
//   ldc = sum over (m,p) of KP(m,p).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SHT_ROUND_UP(x, q) ((((x) + (q) - 1) / (q)) * (q))

namespace sht {

constexpr int kPad = 32;  // row / k padding of every table and data block

// Precision of the Legendre GEMMs.
enum Precision : int {
  kFP32 = 1,    // v_mfma_f32_*_f32, fp32 operands
  kBF16X3 = 2,  // bf16 hi+lo pairs, 3 products (hi*hi + hi*lo + lo*hi)
  kBF16X6 = 3,  // bf16 hi+lo+residual pieces, 6 products (all terms to 2^-16)
};

__host__ __device__ inline int k_count(int T, int m, int p) {
  int len = T + 1 - m - p;
  return len > 0 ? (len + 1) / 2 : 0;
}

__host__ __device__ inline int k_padded(int T, int m, int p) {
  int k = k_count(T, m, p);
  return SHT_ROUND_UP(k > 0 ? k : 1, kPad);
}

}  // namespace sht
