// Register-resident radix-16 rounds of a 256-point (8-bit) NTT pass, shared
// by the standalone NTT kernel and the fused key-switching kernels.
//
// A tile holds 16 sub-transforms of 256 points; sub-transform c is owned by
// the 16 threads with row r = 0..15, each holding 16 points v[m].  Two
// ownerships are used:
//   ROWS  j = r + 16 m   (bits 7..4 of j vary inside a thread)
//   COLS  j = 16 r + m   (bits 3..0 vary inside a thread)
// The forward pass runs its local bits 7..4 in ROWS and 3..0 in COLS; the
// inverse pass runs 0..3 in COLS and 4..7 in ROWS.  Butterflies are the
// reference's exact radix-2 formulas (src/operation/ntt.cpp:46-134,
// :177-272); twiddle addressing is derived in the ntt.hip header comment.
#pragma once

#include "device.h"

namespace fhe {

// Operands are lazy residues in [0, 2q), so every guard sees a non-negative
// value and takes the borrow form.
// Cooley-Tukey: p = mont(w, y); x' = guard(x + p, 2q); y' = guard(x + 2q - p, 2q)
__device__ __forceinline__ void ct_butterfly(int64_t &x, int64_t &y, int64_t w,
                                             int64_t q2, uint64_t q,
                                             uint64_t k) {
  const int64_t p = mont_mult_u(w, y, q, k);
  const int64_t x0 = x;
  x = guard_u(add64(x0, p), q2);
  y = guard_u(sub64(add64(x0, q2), p), q2);
}

// Gentleman-Sande: x' = guard(x + y, 2q); y' = mont(w, guard(x + 2q - y, 2q))
__device__ __forceinline__ void gs_butterfly(int64_t &x, int64_t &y, int64_t w,
                                             int64_t q2, uint64_t q,
                                             uint64_t k) {
  const int64_t x0 = x, y0 = y;
  x = guard_u(add64(x0, y0), q2);
  y = mont_mult_u(w, guard_u(sub64(add64(x0, q2), y0), q2), q, k);
}

// Everything a thread needs to address twiddles in one 8-bit pass: the pass
// covers index bits [b, b+8) of a limb of 2^log_n points; hi is the
// sub-transform's index bits above b+8 and r the thread's row.
struct R16Pass {
  const int64_t *tw; // this prime's twiddle table (forward or inverse)
  int64_t n;
  int64_t hi;
  int log_n;
  int b;
  int r;
  int64_t q2;
  uint64_t q, k;
};

// forward local bits 7..4, ROWS ownership
__device__ __forceinline__ void r16_forward_high(int64_t (&v)[16],
                                                 const R16Pass &s) {
#pragma unroll
  for (int u = 7; u >= 4; u--) {
    const int x = u - 4;
    const int stage = s.log_n - 1 - s.b - u;
    const int64_t tb = (int64_t(1) << stage) - 1 + (s.hi << (7 - u));
#pragma unroll
    for (int m = 0; m < 16; m++) {
      if (m & (1 << x))
        continue;
      ct_butterfly(v[m], v[m + (1 << x)], s.tw[tb + (m >> (u - 3))], s.q2,
                   s.q, s.k);
    }
  }
}

// forward local bits 3..0, COLS ownership
__device__ __forceinline__ void r16_forward_low(int64_t (&v)[16],
                                                const R16Pass &s) {
#pragma unroll
  for (int u = 3; u >= 0; u--) {
    const int stage = s.log_n - 1 - s.b - u;
    const int64_t tb = (int64_t(1) << stage) - 1 + (s.hi << (7 - u)) +
                       (int64_t(s.r) << (3 - u));
#pragma unroll
    for (int m = 0; m < 16; m++) {
      if (m & (1 << u))
        continue;
      ct_butterfly(v[m], v[m + (1 << u)], s.tw[tb + (m >> (u + 1))], s.q2,
                   s.q, s.k);
    }
  }
}

// inverse local bits 0..3, COLS ownership
__device__ __forceinline__ void r16_inverse_low(int64_t (&v)[16],
                                                const R16Pass &s) {
#pragma unroll
  for (int u = 0; u <= 3; u++) {
    const int64_t tb = s.n - (s.n >> (s.b + u)) + (s.hi << (7 - u)) +
                       (int64_t(s.r) << (3 - u));
#pragma unroll
    for (int m = 0; m < 16; m++) {
      if (m & (1 << u))
        continue;
      gs_butterfly(v[m], v[m + (1 << u)], s.tw[tb + (m >> (u + 1))], s.q2,
                   s.q, s.k);
    }
  }
}

// inverse local bits 4..7, ROWS ownership
__device__ __forceinline__ void r16_inverse_high(int64_t (&v)[16],
                                                 const R16Pass &s) {
#pragma unroll
  for (int u = 4; u <= 7; u++) {
    const int x = u - 4;
    const int64_t tb = s.n - (s.n >> (s.b + u)) + (s.hi << (7 - u));
#pragma unroll
    for (int m = 0; m < 16; m++) {
      if (m & (1 << x))
        continue;
      gs_butterfly(v[m], v[m + (1 << x)], s.tw[tb + (m >> (u - 3))], s.q2,
                   s.q, s.k);
    }
  }
}

// LDS exchange buffer: 16 sub-transforms x 256 points, padded so that both
// ownerships and the 16 sub-transforms of a row map to distinct banks.
constexpr int kR16Row = 273;
constexpr int kR16Lds = 16 * kR16Row;

__device__ __forceinline__ int r16_lds(int c, int j) {
  return c * kR16Row + j + (j >> 4);
}

// Re-own the thread's 16 points: ROWS -> COLS (to_cols) or COLS -> ROWS.
// Starts with a barrier so the previous use of the buffer has finished.
__device__ __forceinline__ void r16_exchange(int64_t (&v)[16], int64_t *lds,
                                             int c, int r, bool to_cols) {
  __syncthreads();
#pragma unroll
  for (int m = 0; m < 16; m++)
    lds[r16_lds(c, to_cols ? r + 16 * m : 16 * r + m)] = v[m];
  __syncthreads();
#pragma unroll
  for (int m = 0; m < 16; m++)
    v[m] = lds[r16_lds(c, to_cols ? 16 * r + m : r + 16 * m)];
}

} // namespace fhe
