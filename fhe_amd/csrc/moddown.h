// ModDown of one coefficient, shared by the batched and the uneven-level
// kernels.
#pragma once

#include "device.h"

namespace fhe {

// ModDown (src/ckks/core_key_switch.cpp:110-160) of one key-switched
// polynomial at one coefficient: the special-prime limbs are peeled off from
// the last one down, each subtraction applied to every remaining limb.
// FULL: alpha == MAXA, so that every loop over the special primes has a
// compile-time trip count and the limb loops carry no branches.
template <int MAXA, bool FULL = false> struct ModDown {
  int64_t peeled[MAXA];
  int alpha, limbs, moduli_id;

  __device__ __forceinline__ int specials() const {
    return FULL ? MAXA : alpha;
  }

  __device__ void prepare(const int64_t *a, int64_t n, const DeviceTables &t) {
    int64_t sv[MAXA];
#pragma unroll
    for (int j = 0; j < MAXA; j++)
      sv[j] = j < specials() ? a[int64_t(limbs + j) * n] : 0;
#pragma unroll
    for (int sp = 0; sp < MAXA; sp++) {
      if (sp < specials()) {
        const int r = specials() - 1 - sp;
        int64_t xr = 0;
#pragma unroll
        for (int j = 0; j < MAXA; j++)
          if (j == r)
            xr = sv[j];
        peeled[sp] = xr;
#pragma unroll
        for (int j = 0; j < MAXA; j++) {
          if (j < r) {
            const int p = moduli_id + limbs + j;
            const int64_t q2 = int64_t(t.q2[p]);
            sv[j] = mont_mult(guard(sv[j] + q2 - xr, q2),
                              t.p_inverse[int64_t(sp) * t.prime_count + p],
                              t.q[p], t.k[p]);
          }
        }
      }
    }
  }

  // The canonical key-switched value of ordinary limb i from its
  // accumulator value v >= 0.  The reference's steps
  //   v = mont(guard(v + 2q - p_j, 2q), s_j^-1 R)   (signed), then
  //   make_unsigned_reduce_2q_to_q
  // give the canonical residue of (..((v - p_0) s_0^-1 - p_1) s_1^-1 ..) mod
  // q for the exact peeled integers p_j, whatever representatives the steps
  // carry; here each step is one Shoup product of v + q_top - p_j with the
  // standard-form inverse.  v + q_top - p_j lies in [0, 2^64) because every
  // p_j is a signed Montgomery product with a constant below 2^61, hence in
  // (-2^62, 2^62 + 2^61), and q_top is in [2^63 - q, 2^63).
  __device__ int64_t limb_lazy(int64_t v, int i, const DeviceTables &t) const {
    const int p = moduli_id + i;
    const uint64_t q = t.q[p], top = t.q_top[p];
    uint64_t x = uint64_t(v);
#pragma unroll
    for (int sp = 0; sp < MAXA; sp++)
      if (sp < specials()) {
        const uint64_t *c =
            t.p_inverse_shoup + (int64_t(sp) * t.prime_count + p) * 2;
        x = shoup_mult(x + top - uint64_t(peeled[sp]), c[0], c[1], q);
      }
    return guard_u(int64_t(x), int64_t(q));
  }
};

} // namespace fhe
