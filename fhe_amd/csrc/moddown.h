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
  int64_t neg_peeled[MAXA];  // -p_j through opaque_neg
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
        neg_peeled[sp] = opaque_neg(xr);
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
  // (-2^62, 2^62 + 2^61), and q_top is in [2^63 - q, 2^63).  Every
  // subtraction is an add of a negated value (-p_j, 2^64 - q), one 64-bit
  // add instead of a borrow pair.
  __device__ int64_t limb_lazy(int64_t v, int i, const DeviceTables &t) const {
    const int p = moduli_id + i;
    const uint64_t q = t.q[p], top = t.q_top[p];
    const uint64_t nq = uint64_t(opaque_neg(int64_t(q)));
    uint64_t x = uint64_t(v);
#pragma unroll
    for (int sp = 0; sp < MAXA; sp++)
      if (sp < specials()) {
        const uint64_t *c =
            t.p_inverse_shoup + (int64_t(sp) * t.prime_count + p) * 2;
        x = shoup_mult_neg(x + top + uint64_t(neg_peeled[sp]), c[0], c[1],
                           nq);
      }
    return int64_t(x < q ? x : x + nq);
  }
};

// Limbs per group of the ModDown kernels' limb loops: one thread holds one
// coefficient, so the loads of a group, issued together ahead of its
// arithmetic, are what keeps HBM busy.
constexpr int kMdGroup = 8;

// The relinearized limbs y_i = reduce(d_i + ks_i) (src/ckks/core.cpp:
// 736-737), i in [0, limbs), of one coefficient: a = the key-switched
// accumulator's limbs (limb stride n, the special limbs after the chain
// limbs), d = the addend's limbs when has_d (else any readable limbs, such
// as a's: they are loaded and discarded, so that every load is
// unconditional), o = the output limbs.  A canonical ks_i passes the two
// guards unchanged when there is no addend.  The slots of the last group
// past the end repeat the last limb (its loads, arithmetic and store, which
// writes the same value again), so that a group is one branch-free block.
// The pointers are restrict-qualified: o never overlaps a or the tables,
// and where o is d's own buffer each element is read before the same thread
// writes it.
template <int MAXA, bool FULL>
__device__ __forceinline__ void moddown_relin(const int64_t *__restrict__ a,
                                              const int64_t *__restrict__ d,
                                              bool has_d,
                                              int64_t *__restrict__ o,
                                              int limbs, int moduli_id,
                                              int64_t n,
                                              const DeviceTables &t) {
  ModDown<MAXA, FULL> md;
  md.alpha = t.alpha;
  md.limbs = limbs;
  md.moduli_id = moduli_id;
  md.prepare(a, n, t);
  for (int i0 = 0; i0 < limbs; i0 += kMdGroup) {
    int64_t av[kMdGroup], dv[kMdGroup];
#pragma unroll
    for (int g = 0; g < kMdGroup; g++) {
      const int i = min(i0 + g, limbs - 1);
      av[g] = a[int64_t(i) * n];
      dv[g] = d[int64_t(i) * n];
    }
#pragma unroll
    for (int g = 0; g < kMdGroup; g++) {
      const int i = min(i0 + g, limbs - 1);
      const int p = moduli_id + i;
      const int64_t ks = md.limb_lazy(av[g], i, t);
      const int64_t di = has_d ? dv[g] : 0;
      o[int64_t(i) * n] =
          guard(guard(di + ks, int64_t(t.q2[p])), int64_t(t.q[p]));
    }
  }
}

} // namespace fhe
