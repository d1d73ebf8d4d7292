// Register-resident radix-16 rounds of a 256-point (8-bit) NTT pass, shared
// by the standalone NTT kernel and the fused key-switching kernels.
//
// A tile holds 16 sub-transforms of 256 points; sub-transform c is owned by
// the 16 threads with row r = 0..15, each holding 16 points v[m].  Two
// ownerships are used:
//   ROWS  j = r + 16 m   (bits 7..4 of j vary inside a thread)
//   COLS  j = 16 r + m   (bits 3..0 vary inside a thread)
// The forward pass runs its local bits 7..4 in ROWS and 3..0 in COLS; the
// inverse pass runs 0..3 in COLS and 4..7 in ROWS.  Twiddle addressing is
// derived in the ntt.hip header comment.
//
// Three arithmetics share the loop structure (template parameter A):
//   kExact      the reference's radix-2 formulas with Montgomery twiddles
//               (src/operation/ntt.cpp:46-134, :177-272), so every lazy
//               representative in [0, 2q) matches the reference;
//   kLazySmall  for q < 2^42: Shoup products with standard-form twiddles and
//               no reductions inside the transform.  Forward values grow by
//               less than 2q per stage; inverse sums double per stage and the
//               difference of global stage t is offset by q << (1 + t) (a
//               bound of every input of that stage when the transform's input
//               is below 2q).  Sixteen stages stay below 2^60.
//   kLazyLarge  Harvey's lazy forms for any q < 2^61: forward values in
//               [0, 4q) with one conditional subtraction per butterfly,
//               inverse values in [0, 2q).
// The lazy forms give other representatives of the same residues; they serve
// stages whose results are canonicalized before anything representative-
// dependent reads them.
#pragma once

#include "device.h"

namespace fhe {

enum NttArith : int { kExact = 0, kLazySmall = 1, kLazyLarge = 2 };

// Primes below this bound take the kLazySmall forms.
constexpr uint64_t kLazySmallBound = uint64_t(1) << 42;

// Host side: whether prime p takes the kLazySmall forms.
inline bool t_host_q_small(const DeviceTables &t, int p) {
  return t.host_q[p] < kLazySmallBound;
}

// Exact butterflies with the reference's signed arithmetic.  Operands are
// the reference's representatives: lazy residues in [0, 2q), and at the
// input of a transform possibly negative values (limbs that a larger prime
// rescaled lie in (-q, q)).  The signed product and compares reproduce the
// reference for both; on non-negative operands they equal the unsigned
// forms.  Cooley-Tukey: p = mont(w, y); x' = guard(x + p, 2q);
// y' = guard(x + 2q - p, 2q).  nq2 = -2q, which the guards add in place of
// subtracting 2q (guard_neg); the overloads without it serve callers that
// do not hold it.
__device__ __forceinline__ void ct_butterfly(int64_t &x, int64_t &y, int64_t w,
                                             int64_t q2, int64_t nq2,
                                             uint64_t q, uint64_t k) {
  const int64_t p = mont_mult_cs(w, y, q, k);
  const int64_t x0 = x;
  x = guard_neg(add64(x0, p), q2, nq2);
  y = guard_neg(sub64(add64(x0, q2), p), q2, nq2);
}
__device__ __forceinline__ void ct_butterfly(int64_t &x, int64_t &y, int64_t w,
                                             int64_t q2, uint64_t q,
                                             uint64_t k) {
  ct_butterfly(x, y, w, q2, -q2, q, k);
}

// Gentleman-Sande: x' = guard(x + y, 2q); y' = mont(w, guard(x + 2q - y, 2q))
__device__ __forceinline__ void gs_butterfly(int64_t &x, int64_t &y, int64_t w,
                                             int64_t q2, int64_t nq2,
                                             uint64_t q, uint64_t k) {
  const int64_t x0 = x, y0 = y;
  x = guard_neg(add64(x0, y0), q2, nq2);
  y = mont_mult_cs(w, guard_neg(sub64(add64(x0, q2), y0), q2, nq2), q, k);
}
__device__ __forceinline__ void gs_butterfly(int64_t &x, int64_t &y, int64_t w,
                                             int64_t q2, uint64_t q,
                                             uint64_t k) {
  gs_butterfly(x, y, w, q2, -q2, q, k);
}

// Load-time product with a per-prime constant (mont_enter by R^2 for the
// reference's to_ntt) of a possibly negative input value.  The lazy
// arithmetic takes non-negative operands only: the signed product lies in
// (-q, q), and a negative one is replaced by the residue's representative
// in [0, q).
template <int A>
__device__ __forceinline__ int64_t load_product(int64_t v, int64_t c,
                                                uint64_t q, uint64_t k) {
  int64_t y = mont_mult_cs(c, v, q, k);
  if (A != kExact && y < 0)
    y += int64_t(q);
  return y;
}

// Everything a thread needs to address twiddles in one 8-bit pass: the pass
// covers index bits [b, b+8) of a limb of 2^log_n points; hi is the
// sub-transform's index bits above b+8 and r the thread's row.
struct R16Pass {
  const int64_t *tw;   // Montgomery twiddles of this prime (kExact)
  const uint64_t *ts;  // {w, w'} pairs of this prime (lazy forms)
  int64_t n;
  int64_t hi;
  int log_n;
  int b;
  int r;
  int64_t q2;
  int64_t nq2;  // -q2 through opaque_neg, for the exact arithmetic
  uint64_t q, k;
  uint64_t nq;  // 2^64 - q through opaque_neg, for the lazy forms
};

// One forward butterfly on (x, y) with twiddle entry idx.
template <int A>
__device__ __forceinline__ void fwd_bfly(int64_t &x, int64_t &y,
                                         const R16Pass &s, int64_t idx) {
  if (A == kExact) {
    ct_butterfly(x, y, s.tw[idx], s.q2, s.nq2, s.q, s.k);
    return;
  }
  const uint64_t w = s.ts[2 * idx], wq = s.ts[2 * idx + 1];
  int64_t x0 = x;
  if (A == kLazyLarge)
    x0 = guard_u(x0, s.q2);
  const int64_t t = int64_t(shoup_mult_neg(uint64_t(y), w, wq, s.nq));
  x = add64(x0, t);
  y = sub64(add64(x0, s.q2), t);
}

// One inverse butterfly at global stage t.
template <int A>
__device__ __forceinline__ void inv_bfly(int64_t &x, int64_t &y,
                                         const R16Pass &s, int64_t idx,
                                         int t) {
  if (A == kExact) {
    gs_butterfly(x, y, s.tw[idx], s.q2, s.nq2, s.q, s.k);
    return;
  }
  const uint64_t w = s.ts[2 * idx], wq = s.ts[2 * idx + 1];
  int64_t sum = add64(x, y);
  int64_t off = s.q2;
  if (A == kLazyLarge)
    sum = guard_u(sum, s.q2);
  else
    off = int64_t(s.q << (1 + t));
  const int64_t d = sub64(add64(x, off), y);
  x = sum;
  y = int64_t(shoup_mult_neg(uint64_t(d), w, wq, s.nq));
}

// forward local bits 7..4, ROWS ownership
template <int A = kExact>
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
      fwd_bfly<A>(v[m], v[m + (1 << x)], s, tb + (m >> (u - 3)));
    }
  }
}

// forward local bits 3..0, COLS ownership
template <int A = kExact>
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
      fwd_bfly<A>(v[m], v[m + (1 << u)], s, tb + (m >> (u + 1)));
    }
  }
}

// inverse local bits 0..3, COLS ownership
template <int A = kExact>
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
      inv_bfly<A>(v[m], v[m + (1 << u)], s, tb + (m >> (u + 1)), s.b + u);
    }
  }
}

// inverse local bits 4..7, ROWS ownership
template <int A = kExact>
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
      inv_bfly<A>(v[m], v[m + (1 << x)], s, tb + (m >> (u - 3)), s.b + u);
    }
  }
}

// Bring lazy forward outputs to [0, 2q) (kLazyLarge values are below 4q;
// kLazySmall values are left as they are: their consumers accept them).
template <int A>
__device__ __forceinline__ void r16_forward_settle(int64_t (&v)[16],
                                                   const R16Pass &s) {
  if (A != kLazyLarge)
    return;
#pragma unroll
  for (int m = 0; m < 16; m++)
    v[m] = guard_u(v[m], s.q2);
}

// LDS exchange buffer: 16 sub-transforms x 256 points, padded so that both
// ownerships and the 16 sub-transforms of a row map to distinct banks.
constexpr int kR16Row = 273;
constexpr int kR16Lds = 16 * kR16Row;

__device__ __forceinline__ int r16_lds(int c, int j) {
  return c * kR16Row + j + (j >> 4);
}

// Orders one wavefront's LDS accesses: a wave's LDS instructions execute in
// issue order, so it is enough that the compiler moves no access across.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Re-own the thread's 16 points: ROWS -> COLS (to_cols) or COLS -> ROWS.
// Starts with a barrier so the previous use of the buffer has finished.
// WAVE: the 16 threads sharing c are 16 consecutive threads (c = tid >> 4),
// hence in one wavefront, and a wavefront barrier replaces the block's.
template <bool WAVE = false>
__device__ __forceinline__ void r16_exchange(int64_t (&v)[16], int64_t *lds,
                                             int c, int r, bool to_cols) {
  if (WAVE)
    wave_lds_sync();
  else
    __syncthreads();
#pragma unroll
  for (int m = 0; m < 16; m++)
    lds[r16_lds(c, to_cols ? r + 16 * m : 16 * r + m)] = v[m];
  if (WAVE)
    wave_lds_sync();
  else
    __syncthreads();
#pragma unroll
  for (int m = 0; m < 16; m++)
    v[m] = lds[r16_lds(c, to_cols ? 16 * r + m : r + 16 * m)];
}

// r16_exchange through a buffer of 32-bit words, half the LDS of the 64-bit
// exchange (so that twice as many blocks fit on a CU): the low words go
// through first, then the high words, each between two barriers.
constexpr int kR16LdsWords = kR16Lds;

__device__ __forceinline__ void r16_exchange_words(int64_t (&v)[16],
                                                   uint32_t *lds, int c,
                                                   int r, bool to_cols) {
  uint32_t lo[16];
  __syncthreads();
#pragma unroll
  for (int m = 0; m < 16; m++)
    lds[r16_lds(c, to_cols ? r + 16 * m : 16 * r + m)] = uint32_t(v[m]);
  __syncthreads();
#pragma unroll
  for (int m = 0; m < 16; m++)
    lo[m] = lds[r16_lds(c, to_cols ? 16 * r + m : r + 16 * m)];
  __syncthreads();
#pragma unroll
  for (int m = 0; m < 16; m++)
    lds[r16_lds(c, to_cols ? r + 16 * m : 16 * r + m)] =
        uint32_t(uint64_t(v[m]) >> 32);
  __syncthreads();
#pragma unroll
  for (int m = 0; m < 16; m++)
    v[m] = int64_t(
        (uint64_t(lds[r16_lds(c, to_cols ? 16 * r + m : r + 16 * m)]) << 32) |
        lo[m]);
}

// One register-resident 8-bit pass over one 4096-point tile of one limb:
// the 16 sub-transforms of the tile are owned by 16 threads each (16 points
// per thread) and the eight radix-2 stages run as two radix-16 rounds with
// one LDS exchange between them.  A second exchange is added where the
// global side of the pass is contiguous and the register ownership is not,
// so that every global access is coalesced.  src / dst point at the limb
// (prime p); the pass covers index bits [b, b + 8); COLS is b > 0.  pre
// (forward, per prime) multiplies every loaded value first; post (inverse)
// is the InttPost epilogue; last = the pass completes a forward transform.
// OPAQUE: the pass runs inside a persistent loop (opaque_tid).
template <bool INV, bool COLS, int A, bool OPAQUE = false>
__device__ __forceinline__ void r16_limb_pass(const int64_t *src_limb,
                                              int64_t *dst_limb, int p,
                                              int tile, int b,
                                              const int64_t *pre, int post,
                                              bool last, const DeviceTables &t,
                                              int64_t *lds) {
  const int log_n = t.log_n;
  const int64_t n = int64_t(1) << log_n;
  const int tid = OPAQUE ? opaque_tid() : int(threadIdx.x);
  const int c = COLS ? (tid & 15) : (tid >> 4);
  const int r = COLS ? (tid >> 4) : (tid & 15);
  const int64_t sp = int64_t(tile) * 16 + c;
  const int64_t lo = sp & ((int64_t(1) << b) - 1);
  const int64_t hi = sp >> b;
  const int64_t base = (hi << (b + 8)) + lo;
  const int64_t *src = src_limb + base;
  int64_t *dst = dst_limb + base;

  R16Pass s;
  s.tw = (INV ? t.root_inverse : t.root) + int64_t(p) * n;
  s.ts = (INV ? t.root_inverse_shoup : t.root_shoup) + int64_t(p) * n * 2;
  s.n = n;
  s.hi = hi;
  s.log_n = log_n;
  s.b = b;
  s.r = r;
  s.q = t.q[p];
  s.k = t.k[p];
  s.q2 = int64_t(t.q2[p]);
  s.nq2 = A == kExact ? opaque_neg(s.q2) : 0;  // exact guards only
  s.nq = A == kExact ? 0 : uint64_t(opaque_neg(int64_t(s.q)));

  int64_t v[16];
  // load in the ownership of the first round; the contiguous inverse pass
  // loads coalesced ROWS and re-owns through LDS
  const bool cols_load = INV && COLS;
#pragma unroll
  for (int m = 0; m < 16; m++)
    v[m] = src[int64_t(cols_load ? 16 * r + m : r + 16 * m) << b];
  if (pre) {
    const int64_t cst = pre[p];
#pragma unroll
    for (int m = 0; m < 16; m++)
      v[m] = load_product<A>(v[m], cst, s.q, s.k);
  }
  if (INV && !COLS)
    r16_exchange<!COLS>(v, lds, c, r, true);

  if (!INV) {
    r16_forward_high<A>(v, s);
    r16_exchange<!COLS>(v, lds, c, r, true);
    r16_forward_low<A>(v, s);
    if (last)
      r16_forward_settle<A>(v, s);
    if (!COLS)
      r16_exchange<!COLS>(v, lds, c, r, false); // coalesced contiguous store
  } else {
    r16_inverse_low<A>(v, s);
    r16_exchange<!COLS>(v, lds, c, r, false);
    r16_inverse_high<A>(v, s);
    if (post != kInttRaw) {
      const int64_t ninv = t.ninv[p];
      if (A == kExact) {
#pragma unroll
        for (int m = 0; m < 16; m++) {
          // only the exact arithmetic carries negative values
          const int64_t y =
              mont_exit(mont_mult_cs(ninv, v[m], s.q, s.k), s.q, s.k);
          v[m] = post == kInttExitReduce ? guard(y, int64_t(s.q)) : y;
        }
      } else {
        // the lazy forms need only the residue: one product with
        // N^-1 R^-1 (the exit folded into the constant) gives it below
        // v q / 2^62 + q < 2q (v < 2^60 for kLazySmall, v < 2q for
        // kLazyLarge), which the reduce brings to the canonical value
        const int64_t ninv_exit = mont_exit(ninv, s.q, s.k);
#pragma unroll
        for (int m = 0; m < 16; m++) {
          const int64_t y = mont_mult_u(v[m], ninv_exit, s.q, s.k);
          v[m] = post == kInttExitReduce ? guard_u(y, int64_t(s.q)) : y;
        }
      }
    }
  }
  // store: ROWS ownership except the strided forward pass (COLS)
  const bool cols_store = !INV && COLS;
#pragma unroll
  for (int m = 0; m < 16; m++)
    dst[int64_t(cols_store ? 16 * r + m : r + 16 * m) << b] = v[m];
}

} // namespace fhe
