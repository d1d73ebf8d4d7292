// Per-coefficient ModUp arithmetic shared by the standalone and the fused
// (ModUp + first forward NTT pass) kernels: Garner digits of one partition
// (rns_to_mrs, src/ckks/core_key_switch.cpp:21-56) and the exact basis
// extension of one limb (mod_up, :92-108), with the reference's operation
// order and guards.
#pragma once

#include "device.h"
#include "params.h"

namespace fhe {

// Montgomery product of a value a with a table constant b for ModUp: the
// general signed form for the signed partitions (whose values may exceed
// 2^62), else the fast form for values in [-2^62, 2^62) (inputs rescaled by
// a larger prime can be negative).
template <bool SIGNED>
__device__ __forceinline__ int64_t modup_mult(int64_t a, int64_t b,
                                              uint64_t q, uint64_t k) {
  return SIGNED ? mont_mult(a, b, q, k) : mont_mult_cs(b, a, q, k);
}

// d[j] = the partition's Garner digits from its limb values x[j] (limb
// start + j at moduli_id); cst points at the partition's constants.
template <int MAXZ, bool SIGNED>
__device__ __forceinline__ void garner_digits(const int64_t (&x)[MAXZ],
                                              int64_t (&d)[MAXZ], int size,
                                              int pbase, const int64_t *cst,
                                              const DeviceTables &t) {
#pragma unroll
  for (int j = 0; j < MAXZ; j++)
    d[j] = x[0];
#pragma unroll
  for (int j = 1; j < MAXZ; j++) {
    if (j < size) {
      const int pj = pbase + j;
      const int64_t m = int64_t(t.q[pj]), m2 = int64_t(t.q2[pj]);
      const int64_t diff = guard_fast(sub64(add64(x[j], m2), d[j]), m2);
      d[j] = guard_fast(modup_mult<SIGNED>(diff,
                                           cst[CkksParams::kKsInvOffset + j],
                                           t.q[pj], t.k[pj]),
                        m);
#pragma unroll
      for (int tt = j + 1; tt < MAXZ; tt++) {
        if (tt < size) {
          const int pt = pbase + tt;
          const int64_t prod = modup_mult<SIGNED>(
              d[j], cst[CkksParams::kKsTileOffset + j * CkksParams::kMaxAlpha +
                        tt],
              t.q[pt], t.k[pt]);
          d[tt] = guard_fast(guard_fast(add64(prod, d[tt]), int64_t(t.q2[pt])),
                             int64_t(t.q[pt]));
        }
      }
    }
  }
}

// Extended limb i (prime p) from the digits: sum_j d_j * (prod q_<j R^2).
template <int MAXZ, bool SIGNED>
__device__ __forceinline__ int64_t extend_limb(const int64_t (&d)[MAXZ],
                                               int size, int i, int ext_limbs,
                                               const int64_t *ec, uint64_t q,
                                               uint64_t k, int64_t q2) {
  int64_t y = modup_mult<SIGNED>(d[0], ec[i], q, k);
#pragma unroll
  for (int j = 1; j < MAXZ; j++)
    if (j < size)
      y = guard_fast(
          add64(modup_mult<SIGNED>(d[j], ec[j * ext_limbs + i], q, k), y), q2);
  // a negative digit can leave y in (-q, 0): the extended limbs only reach
  // the (unsigned) transforms and the key products, whose results are
  // canonicalized before use, so the non-negative representative y + 2q
  // serves in its place
  if (y < 0)
    y += q2;
  return y;
}

// The same extended limb as any representative of its residue, for the
// lazy key switch (the extended limbs only reach ModDown through the
// canonicalizing INTT): Shoup products with the standard-form constants
// (pairs {c, floor(c 2^64 / q)} at es + 2 (j ext + i)).  Primes below 2^42
// keep the plain sum (below 2 * size * q); larger primes keep the sum below
// 2q, as the Harvey forward transform expects.  nq = 2^64 - q, opaque to the
// compiler (opaque_neg), for the products and the subtraction of 2q.
template <int MAXZ>
__device__ __forceinline__ int64_t extend_limb_lazy(const int64_t (&d)[MAXZ],
                                                    int size, int i,
                                                    int ext_limbs,
                                                    const uint64_t *es,
                                                    uint64_t q, uint64_t nq,
                                                    bool small) {
  // the first digit is the input value itself, negative for limbs that a
  // larger prime rescaled: it enters as d + big, big a multiple of q in
  // [2^61, 2^62) above every |d| (the same residue mod q).  The other
  // digits of a partition on this path (no signed partitions) are
  // non-negative: each is a Montgomery product of a non-negative difference.
  const uint64_t big = q << (__builtin_clzll(q) - 2);
  const uint64_t *c0 = es + 2 * i;
  uint64_t y = shoup_mult_neg(uint64_t(d[0]) + (d[0] < 0 ? big : 0), c0[0],
                              c0[1], nq);
#pragma unroll
  for (int j = 1; j < MAXZ; j++)
    if (j < size) {
      const uint64_t *c = es + 2 * (j * ext_limbs + i);
      y += shoup_mult_neg(uint64_t(d[j]), c[0], c[1], nq);
      if (!small && y >= 2 * q)
        y += 2 * nq;
    }
  return int64_t(y);
}

} // namespace fhe
