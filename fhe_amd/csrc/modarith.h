// Montgomery modular arithmetic with the reference's exact scalar semantics.
//
// Radix R = 2^62, k = -q^-1 mod 2^62 and a product that is
//   r = (x + ((x * k) mod 2^62) * q) >> 62,   x = a * b   (128-bit),
// with operands sign-extended to 128 bits, as in
// src/montgomery/unit.h:13-31.  Every helper here returns bit-identical
// int64 values to that code for every int64 input, including the negative
// intermediates that ModDown and rescale produce, so device results can be
// compared limb for limb with the reference.
//
// gfx950 has no 64x64-bit multiplier: the device path builds the 128-bit
// products from four 32x32->64 multiply-adds (v_mad_u64_u32).
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define FHE_HD __host__ __device__ __forceinline__
#else
#define FHE_HD inline
#endif

namespace fhe {

constexpr int kMontBits = 62;
constexpr uint64_t kMontMask = (uint64_t(1) << kMontBits) - 1;
constexpr int64_t kMontR = int64_t(1) << kMontBits;

struct U128 {
  uint64_t lo, hi;
};

// Full 64x64 -> 128-bit unsigned product from 32-bit halves.
FHE_HD U128 mul_wide(uint64_t a, uint64_t b) {
  const uint32_t a0 = uint32_t(a), a1 = uint32_t(a >> 32);
  const uint32_t b0 = uint32_t(b), b1 = uint32_t(b >> 32);
  const uint64_t p00 = uint64_t(a0) * b0;
  const uint64_t p01 = uint64_t(a0) * b1 + (p00 >> 32);
  const uint64_t p10 = uint64_t(a1) * b0 + uint32_t(p01);
  const uint64_t p11 = uint64_t(a1) * b1 + (p01 >> 32) + (p10 >> 32);
  return {(p10 << 32) | uint32_t(p00), p11};
}

// (x + s*q) >> 62 truncated to 64 bits, where the caller has already
// formed the 128-bit x; s = (x.lo * k) mod 2^62.
FHE_HD uint64_t mont_reduce(U128 x, uint64_t q, uint64_t k) {
  const uint64_t s = (x.lo * k) & kMontMask;
  const U128 sq = mul_wide(s, q);
  const uint64_t lo = x.lo + sq.lo;
  const uint64_t hi = x.hi + sq.hi + (lo < x.lo ? 1 : 0);
  return (hi << 2) | (lo >> kMontBits);
}

// 32-bit add with carry in/out (v_add_co_u32 / v_addc_co_u32 on gfx950), for
// the word sums of 128-bit products, where a carry out of the low half is
// needed.
FHE_HD uint32_t add_carry(uint32_t a, uint32_t b, uint32_t cin,
                          uint32_t &cout) {
#if defined(__clang__)
  unsigned int co;
  const uint32_t r = __builtin_addc(a, b, cin, &co);
  cout = co;
  return r;
#else
  const uint64_t s = uint64_t(a) + b + cin;
  cout = uint32_t(s >> 32);
  return uint32_t(s);
#endif
}

// 128-bit product of a, b < 2^63 as four 32-bit words.  The two cross
// products are summed inside one v_mad_u64_u32 chain: a1*b0 + a0*b1 < 2^64
// because both high halves are below 2^31.
struct Words4 {
  uint32_t w0, w1, w2, w3;
};
FHE_HD Words4 mul_words(uint64_t a, uint64_t b) {
  const uint32_t a0 = uint32_t(a), a1 = uint32_t(a >> 32);
  const uint32_t b0 = uint32_t(b), b1 = uint32_t(b >> 32);
  const uint64_t p00 = uint64_t(a0) * b0;
  const uint64_t mid = uint64_t(a1) * b0 + uint64_t(a0) * b1;
  const uint64_t p11 = uint64_t(a1) * b1;
  uint32_t c;
  Words4 x;
  x.w0 = uint32_t(p00);
  x.w1 = add_carry(uint32_t(p00 >> 32), uint32_t(mid), 0, c);
  x.w2 = add_carry(uint32_t(p11), uint32_t(mid >> 32), c, c);
  x.w3 = uint32_t(p11 >> 32) + c;
  return x;
}

// Montgomery product of two NON-NEGATIVE operands below 2^63 (lazy residues
// in [0, 2q) and table constants in [0, q)).  Same value as mont_mult.
// On the device it uses no carry chain (a carry read through VCC costs wait
// states): x = a b and y = s q come from multiply-adds as hi 2^64 + lo, and
// since lo(x) + lo(y) = 0 mod 2^62 their low 62 bits sum to 0 (both zero)
// or exactly 2^62 (otherwise), so
//   (x + y) / 2^62 = 4 (hi(x) + hi(y)) + lo(x) / 2^62 + lo(y) / 2^62
//                    + [lo(x) mod 2^62 != 0].
FHE_HD int64_t mont_mult_u(int64_t a, int64_t b, uint64_t q, uint64_t k) {
#if defined(__HIP_DEVICE_COMPILE__)
  const uint32_t a0 = uint32_t(a), a1 = uint32_t(uint64_t(a) >> 32);
  const uint32_t b0 = uint32_t(b), b1 = uint32_t(uint64_t(b) >> 32);
  // x = a b: x0, x1 its low words; u >> 32 carries into the high half
  // (a1 b0 + a0 b1 + 2^32 < 2^64 for a, b < 2^63)
  const uint64_t p00 = uint64_t(a0) * b0;
  const uint64_t u = uint64_t(a0) * b1 + (uint64_t(a1) * b0 + (p00 >> 32));
  const uint32_t x0 = uint32_t(p00), x1 = uint32_t(u);
  // s = (x mod 2^64) * k mod 2^62
  const uint32_t k0 = uint32_t(k), k1 = uint32_t(k >> 32);
  const uint64_t t = uint64_t(x0) * k0;
  const uint32_t s0 = uint32_t(t);
  const uint32_t s1 =
      (uint32_t(t >> 32) + x0 * k1 + x1 * k0) & ((1u << 30) - 1);
  // y = s q: its second word is uint32_t(v)
  const uint32_t q0 = uint32_t(q), q1 = uint32_t(q >> 32);
  const uint64_t r00 = uint64_t(s0) * q0;
  const uint64_t v = uint64_t(s0) * q1 + (uint64_t(s1) * q0 + (r00 >> 32));
  // hi(x) + hi(y) (below 2^63)
  const uint64_t hi =
      uint64_t(a1) * b1 + uint64_t(s1) * q1 + (u >> 32) + (v >> 32);
  const uint32_t low = x0 | (x1 << 2); // lo(x) mod 2^62, bits moved
  const uint32_t nonzero = low < 1u ? low : 1u;
  const uint32_t small = (x1 >> 30) + (uint32_t(v) >> 30) + nonzero;
  return int64_t((hi << 2) + small);
#else
  return int64_t(mont_reduce(mul_wide(uint64_t(a), uint64_t(b)), q, k));
#endif
}

// 64-bit add: one v_lshl_add_u64 on gfx950 (a single full-rate
// instruction, measured at twice the throughput of a v_add_co_u32 /
// v_addc_co_u32 pair, whose carry read through VCC also costs two wait
// states when the pair issues back to back).
FHE_HD int64_t add64(int64_t a, int64_t b) {
  return int64_t(uint64_t(a) + uint64_t(b));
}

// 64-bit subtract as two 32-bit borrow ops on the device.
FHE_HD int64_t sub64(int64_t a, int64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  unsigned int bo;
  const uint32_t lo = __builtin_subc(uint32_t(a), uint32_t(b), 0u, &bo);
  const uint32_t hi = __builtin_subc(uint32_t(uint64_t(a) >> 32),
                                     uint32_t(uint64_t(b) >> 32), bo, &bo);
  return int64_t((uint64_t(hi) << 32) | lo);
#else
  return int64_t(uint64_t(a) - uint64_t(b));
#endif
}

// guard() for |x|, |q| < 2^62 (no signed overflow in x - q): the sign of
// x - q decides, so a 64-bit subtract and a select replace the compare.
FHE_HD int64_t guard_fast(int64_t x, int64_t q) {
  const int64_t d = sub64(x, q);
  return int32_t(uint64_t(d) >> 32) < 0 ? x : d;
}

// guard() for NON-NEGATIVE x and q (an unsigned compare).  Written as a
// plain compare-and-select: on gfx950 this measured faster inside the
// butterflies than a borrow-based form.
FHE_HD int64_t guard_u(int64_t x, int64_t q) {
  return uint64_t(x) < uint64_t(q) ? x : x - q;
}

// Montgomery product of arbitrary int64 operands: the 128-bit product of the
// sign-extended values, i.e. the unsigned product corrected in its high word.
FHE_HD int64_t mont_mult(int64_t a, int64_t b, uint64_t q, uint64_t k) {
  U128 x = mul_wide(uint64_t(a), uint64_t(b));
  if (a < 0)
    x.hi -= uint64_t(b);
  if (b < 0)
    x.hi -= uint64_t(a);
  return int64_t(mont_reduce(x, q, k));
}

// mont_mult(c, b) for a constant c in [0, 2^62) and a SIGNED operand b in
// [-2^62, 2^62) (the values the transforms and ModUp see: residues, lazy
// residues, and inputs in (-q, q) with q < 2^61), through the fast unsigned
// product: b' = b + 2^62 is non-negative and below 2^63, and
// mont(c, b') = mont(c, b) + c exactly, because the added c * 2^62 is a
// multiple of the radix (it leaves (x * k) mod 2^62 unchanged and adds
// exactly c to the shifted sum).
FHE_HD int64_t mont_mult_cs(int64_t c, int64_t b, uint64_t q, uint64_t k) {
  return sub64(mont_mult_u(c, add64(b, kMontR), q, k), c);
}

// High 64 bits of a 64 x 64-bit product from 32-bit multiply-adds.
FHE_HD uint64_t mulhi64(uint64_t a, uint64_t b) {
  const uint32_t a0 = uint32_t(a), a1 = uint32_t(a >> 32);
  const uint32_t b0 = uint32_t(b), b1 = uint32_t(b >> 32);
  const uint64_t ll = uint64_t(a0) * b0;
  const uint64_t lh = uint64_t(a0) * b1 + (ll >> 32);
  const uint64_t hl = uint64_t(a1) * b0 + uint32_t(lh);
  return uint64_t(a1) * b1 + (lh >> 32) + (hl >> 32);
}

// Shoup product (a representative of w y mod q in [0, 2q)) for any
// y < 2^64, w < q and wq = floor(w 2^64 / q).
FHE_HD uint64_t shoup_mult(uint64_t y, uint64_t w, uint64_t wq, uint64_t q) {
  return w * y - mulhi64(wq, y) * q;
}

// shoup_mult given nq = 2^64 - q: the same value as w y + h nq (mod 2^64),
// a sum the multiply-adds absorb, where w y - h q ends in a borrow pair.
// The caller keeps nq opaque to the compiler (opaque_neg), which would
// otherwise fold the product with -q back into a subtraction.
FHE_HD uint64_t shoup_mult_neg(uint64_t y, uint64_t w, uint64_t wq,
                               uint64_t nq) {
  return w * y + mulhi64(wq, y) * nq;
}


// Montgomery exit x * R^-1 (src/montgomery/unit.h:25-31): the single-word
// operand is sign-extended to 128 bits.
FHE_HD int64_t mont_exit(int64_t x, uint64_t q, uint64_t k) {
  const U128 xx = {uint64_t(x), x < 0 ? ~uint64_t(0) : uint64_t(0)};
  return int64_t(mont_reduce(xx, q, k));
}

// Conditional subtraction (src/montgomery/unit.h:13), signed compare.
FHE_HD int64_t guard(int64_t x, int64_t q) { return (x < q) ? x : x - q; }

// guard(x, q) given nq = -q: the subtraction is one 64-bit add
// (v_lshl_add_u64) instead of a borrow pair.  The caller keeps nq opaque to
// the compiler (opaque_neg), which would otherwise fold x + (-q) back into
// x - q.
FHE_HD int64_t guard_neg(int64_t x, int64_t q, int64_t nq) {
  return (x < q) ? x : add64(x, nq);
}

} // namespace fhe
