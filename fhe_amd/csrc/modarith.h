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

// Montgomery product of two NON-NEGATIVE operands (the common case: lazy
// residues in [0, 2q) and table constants in [0, q)).
FHE_HD int64_t mont_mult_u(int64_t a, int64_t b, uint64_t q, uint64_t k) {
  return int64_t(mont_reduce(mul_wide(uint64_t(a), uint64_t(b)), q, k));
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

// Montgomery exit x * R^-1 (src/montgomery/unit.h:25-31): the single-word
// operand is sign-extended to 128 bits.
FHE_HD int64_t mont_exit(int64_t x, uint64_t q, uint64_t k) {
  const U128 xx = {uint64_t(x), x < 0 ? ~uint64_t(0) : uint64_t(0)};
  return int64_t(mont_reduce(xx, q, k));
}

// Conditional subtraction (src/montgomery/unit.h:13), signed compare.
FHE_HD int64_t guard(int64_t x, int64_t q) { return (x < q) ? x : x - q; }

} // namespace fhe
