// Work items of the fused key-switching kernels for N = 2^16, addressed
// through small per-item views so that the batched kernels (one moduli_id
// for the whole batch) and the uneven-level kernels (a moduli_id per
// ciphertext) share one body.
//
// ks_item_tile: the key inner product of one 4096-coefficient tile of one
// extended limb of one ciphertext (src/ckks/core_key_switch.cpp:293-358):
// for every partition in turn, the low-bit forward NTT pass of the extended
// polynomial (the high-bit pass ran before), the products with both key
// halves and their accumulation in registers; then the low-bit pass of the
// inverse NTT on both accumulators before a single write.
//
// ext_item_tile: one extended limb of one partition formed from the Garner
// digits (src/ckks/core_key_switch.cpp:92-108) with the first (high-bit)
// forward NTT pass fused before its only write.
#pragma once

#include "kernels.h"
#include "modup.h"
#include "ntt_r16.h"

namespace fhe {

// Arguments of the batched key inner product (one moduli_id for the batch).
struct KsArgs {
  const int64_t *ext;  // [count][parts][ext_limbs][N], forward pass 1 done
  const int64_t *key;  // [key_polys][key_chain][N]
  int64_t *acc;        // [count][2][ext_limbs][N], inverse pass 1 done
  int count;
  int parts;
  int ext_limbs;
  int moduli_id;
  int key_chain;
  int key_offset;      // first key limb aligned with the ciphertext's limb 0
  int key_partition0;  // key partition index of the ciphertext's partition 0
  int swizzle;
  int lazy;            // lazy transforms (the ext pass before was lazy too)
  // NTT form of the key-switched polynomial, when the caller has it: limb i
  // of ciphertext c at own + c * own_stride + i * N.  The extended limbs of a
  // partition's own primes are that polynomial's limbs, so their transforms
  // are taken from here instead of the extended polynomial (any
  // representative serves: the key products are canonicalized later).
  const int64_t *own;
  int64_t own_stride;
  int start[kMaxFusedParts];
  int size[kMaxFusedParts];
  int lazy_acc;        // small primes: accumulate without guards
};


// The key inner product reaches its data through an addressing policy
// `Addr` (one per kernel family) with the members
//   int parts, ext_limbs, moduli_id, key_chain, lazy_acc;
//   bool is_own(pp, i)          limb i of partition pp is the source's own
//   const int64_t *own(i, off)  that limb's NTT-form tile (thread row 0)
//   const int64_t *ext(pp, i, off)  the extended limb's tile
//   const int64_t *key_tile(pp, i, off)  key half 0 of partition pp (half 1
//                               is key_chain limbs further)
//   int64_t *acc(i, off)        accumulator half 0 (half 1 is ext_limbs
//                               limbs further)
// A partition's own extended limbs are the NTT-form source's limbs, so their
// transforms are taken from there (any representative serves: the key
// products are canonicalized later).

// Sum of key products of one small prime kept as three 32-bit words.  With
// q < 2^42 the lazy forward outputs are below 2q (16 + size) <= 64q and the
// key values below 2q, so each product is below 128 q^2 < 2^91 and a sum of
// at most kMaxWideParts = 16 of them below 2^95.  One Montgomery reduction
// of the sum S returns (S + s q) / 2^62 < S / 2^62 + q < q (parts 128 q /
// 2^62) + q < 2q: the bound the inverse transform's inputs need.
struct Acc96 {
  uint32_t w0, w1, w2;
};
constexpr int kMaxWideParts = 16;

// 96-bit product of a < 2^48 and b < 2^43 (so a1 b1 < 2^27 and the product
// has no fourth word).
__device__ __forceinline__ Words4 mul_words96(uint64_t a, uint64_t b) {
  return mul_words(a, b);
}

__device__ __forceinline__ void acc96_set(Acc96 &s, uint64_t a, uint64_t b) {
  const Words4 x = mul_words96(a, b);
  s = {x.w0, x.w1, x.w2};
}

__device__ __forceinline__ void acc96_add(Acc96 &s, uint64_t a, uint64_t b) {
  const Words4 x = mul_words96(a, b);
  uint32_t c;
  s.w0 = add_carry(s.w0, x.w0, 0, c);
  s.w1 = add_carry(s.w1, x.w1, c, c);
  s.w2 = s.w2 + x.w2 + c;
}

// Montgomery reduction (S + ((S k) mod 2^62) q) >> 62 of a 96-bit sum.
__device__ __forceinline__ int64_t acc96_reduce(const Acc96 &x, uint64_t q,
                                                uint64_t k) {
  const uint32_t k0 = uint32_t(k), k1 = uint32_t(k >> 32);
  const uint64_t t = uint64_t(x.w0) * k0;
  const uint32_t s0 = uint32_t(t);
  const uint32_t s1 =
      (uint32_t(t >> 32) + x.w0 * k1 + x.w1 * k0) & ((1u << 30) - 1);
  const Words4 y = mul_words((uint64_t(s1) << 32) | s0, q);
  uint32_t c;
  (void)add_carry(x.w0, y.w0, 0, c);
  const uint32_t z1 = add_carry(x.w1, y.w1, c, c);
  const uint32_t z2 = add_carry(x.w2, y.w2, c, c);
  const uint32_t z3 = y.w3 + c;
  const uint32_t lo = __builtin_amdgcn_alignbit(z2, z1, 30);
  const uint32_t hi = __builtin_amdgcn_alignbit(z3, z2, 30);
  return int64_t((uint64_t(hi) << 32) | lo);
}

// The (extended limb i, tile) work item of one ciphertext with arithmetic A
// for the two transforms.  The key products are Montgomery products of
// operands whose product stays below q * 2^62, accumulated with the
// reference's guards; on small primes, lazy_acc sums the reduced products
// and brings the sum under 2q once, and wide_acc sums the unreduced 96-bit
// products and reduces each accumulator once (Acc96).  Every form gives the
// same residues, and the accumulators are canonicalized after the inverse
// transform.
// OPAQUE: the tile runs inside a persistent loop (opaque_tid).
template <int A, bool PF, bool WIDE, class Addr, bool OPAQUE = false>
__device__ __forceinline__ void ks_item_tile(const Addr &a,
                                             const DeviceTables &t,
                                             int64_t *lds, int tile, int i) {
  const int64_t n = int64_t(1) << t.log_n;
  const int p = a.moduli_id + i;
  const int tid = OPAQUE ? opaque_tid() : int(threadIdx.x);
  const int r = tid & 15;
  const int c = tid >> 4;
  const int64_t sp = int64_t(tile) * 16 + c; // 256-point block in the limb
  const int64_t off = sp << 8;

  R16Pass f;
  f.tw = t.root + int64_t(p) * n;
  f.ts = t.root_shoup + int64_t(p) * n * 2;
  f.n = n;
  f.hi = sp;
  f.log_n = t.log_n;
  f.b = 0;
  f.r = r;
  f.q = t.q[p];
  f.k = t.k[p];
  f.q2 = int64_t(t.q2[p]);
  f.nq2 = A == kExact ? opaque_neg(f.q2) : 0;  // exact guards only
  f.nq = A == kExact ? 0 : uint64_t(opaque_neg(int64_t(f.q)));
  const int64_t qq = int64_t(f.q);
  // 96-bit sums reduced once: small primes, within the partition bound
  // (WIDE selects the code; the kernel chooses it only for such items)
  constexpr bool wide = WIDE && A == kLazySmall;

  // block-uniform: every thread of the block has the same (i, pp)
  auto is_own = [&](int pp) { return a.is_own(pp, i); };
  auto ext_tile = [&](int pp) { return a.ext(pp, i, off); };
  int64_t acc0[16], acc1[16], v[16];
  Acc96 w0[wide ? 16 : 1], w1[wide ? 16 : 1];
  // with PF the next partition's extended tile is loaded while the current
  // one is transformed and multiplied
  int64_t nx[PF ? 16 : 1];
  if (PF && !is_own(0)) {
    const int64_t *src = ext_tile(0);
#pragma unroll
    for (int m = 0; m < 16; m++)
      nx[m] = src[r + 16 * m];
  }
  for (int pp = 0; pp < a.parts; pp++) {
    const bool own = is_own(pp);
    if (own) {
      const int64_t *src = a.own(i, off) + 16 * r;
#pragma unroll
      for (int m = 0; m < 16; m++)
        v[m] = src[m];
    } else if (PF) {
#pragma unroll
      for (int m = 0; m < 16; m++)
        v[m] = nx[m];
    } else {
      const int64_t *src = ext_tile(pp);
#pragma unroll
      for (int m = 0; m < 16; m++)
        v[m] = src[r + 16 * m];
    }
    if (PF && pp + 1 < a.parts && !is_own(pp + 1)) {
      const int64_t *src = ext_tile(pp + 1);
#pragma unroll
      for (int m = 0; m < 16; m++)
        nx[m] = src[r + 16 * m];
    }
    if (!own) {
      r16_forward_high<A>(v, f);
      r16_exchange<true>(v, lds, c, r, true);
      r16_forward_low<A>(v, f);
      r16_forward_settle<A>(v, f);
    }
    // v[m] is now the NTT value at coefficient off + 16 r + m
    const int64_t *k0 = a.key_tile(pp, i, off) + 16 * r;
    const int64_t *k1 = k0 + int64_t(a.key_chain) * n;
    if constexpr (wide) {
#pragma unroll
      for (int m = 0; m < 16; m++) {
        if (pp == 0) {
          acc96_set(w0[m], uint64_t(v[m]), uint64_t(k0[m]));
          acc96_set(w1[m], uint64_t(v[m]), uint64_t(k1[m]));
        } else {
          acc96_add(w0[m], uint64_t(v[m]), uint64_t(k0[m]));
          acc96_add(w1[m], uint64_t(v[m]), uint64_t(k1[m]));
        }
      }
      continue;
    }
    if (pp == 0) {
#pragma unroll
      for (int m = 0; m < 16; m++) {
        acc0[m] = mont_mult_u(v[m], k0[m], f.q, f.k);
        acc1[m] = mont_mult_u(v[m], k1[m], f.q, f.k);
      }
    } else if (A == kLazySmall && a.lazy_acc) {
      // q < 2^42: the sum of the parts' products (each below 2q) stays far
      // below 2^63; it is brought back under 2q once, after the loop
#pragma unroll
      for (int m = 0; m < 16; m++) {
        acc0[m] = add64(mont_mult_u(v[m], k0[m], f.q, f.k), acc0[m]);
        acc1[m] = add64(mont_mult_u(v[m], k1[m], f.q, f.k), acc1[m]);
      }
    } else if (A == kLazyLarge && a.lazy_acc) {
      // sums kept below 2q (one conditional subtraction of 2q): the lazy
      // inverse transform that follows takes inputs below 2q
#pragma unroll
      for (int m = 0; m < 16; m++) {
        acc0[m] = guard_u(add64(mont_mult_u(v[m], k0[m], f.q, f.k), acc0[m]),
                          f.q2);
        acc1[m] = guard_u(add64(mont_mult_u(v[m], k1[m], f.q, f.k), acc1[m]),
                          f.q2);
      }
    } else {
#pragma unroll
      for (int m = 0; m < 16; m++) {
        acc0[m] = guard_u(
            guard_u(add64(mont_mult_u(v[m], k0[m], f.q, f.k), acc0[m]), f.q2),
            qq);
        acc1[m] = guard_u(
            guard_u(add64(mont_mult_u(v[m], k1[m], f.q, f.k), acc1[m]), f.q2),
            qq);
      }
    }
  }

  if constexpr (wide) {
#pragma unroll
    for (int m = 0; m < 16; m++) {
      acc0[m] = acc96_reduce(w0[m], f.q, f.k);
      acc1[m] = acc96_reduce(w1[m], f.q, f.k);
    }
  } else if (A == kLazySmall && a.lazy_acc && a.parts > 1) {
    // acc < 2q * parts: halving guards from 2q * 2^ceil(log2(parts) - 1)
    int64_t bound = f.q2;
    while (bound < f.q2 * a.parts)
      bound <<= 1;
    for (bound >>= 1; bound >= f.q2; bound >>= 1) {
#pragma unroll
      for (int m = 0; m < 16; m++) {
        acc0[m] = guard_u(acc0[m], bound);
        acc1[m] = guard_u(acc1[m], bound);
      }
    }
  }
  R16Pass g = f;
  g.tw = t.root_inverse + int64_t(p) * n;
  g.ts = t.root_inverse_shoup + int64_t(p) * n * 2;
  int64_t *d0 = a.acc(i, off);
  int64_t *d1 = d0 + int64_t(a.ext_limbs) * n;
  r16_inverse_low<A>(acc0, g);
  r16_exchange<true>(acc0, lds, c, r, false);
  r16_inverse_high<A>(acc0, g);
#pragma unroll
  for (int m = 0; m < 16; m++)
    d0[r + 16 * m] = acc0[m];
  r16_inverse_low<A>(acc1, g);
  r16_exchange<true>(acc1, lds, c, r, false);
  r16_inverse_high<A>(acc1, g);
#pragma unroll
  for (int m = 0; m < 16; m++)
    d1[r + 16 * m] = acc1[m];
}

// Addressing of ciphertext ct of a batch at one moduli_id (KsArgs).
struct KsBatchAddr {
  const KsArgs &a;
  int ct, log_n;
  int parts, ext_limbs, moduli_id, key_chain, lazy_acc;
  __device__ KsBatchAddr(const KsArgs &args, int c, int ln)
      : a(args), ct(c), log_n(ln), parts(args.parts),
        ext_limbs(args.ext_limbs), moduli_id(args.moduli_id),
        key_chain(args.key_chain), lazy_acc(args.lazy_acc) {}
  __device__ bool is_own(int pp, int i) const {
    return a.own != nullptr && i >= a.start[pp] && i < a.start[pp] + a.size[pp];
  }
  __device__ const int64_t *own(int i, int64_t off) const {
    return a.own + int64_t(ct) * a.own_stride + (int64_t(i) << log_n) + off;
  }
  __device__ const int64_t *ext(int pp, int i, int64_t off) const {
    return a.ext + (((int64_t(ct) * a.parts + pp) * a.ext_limbs + i) << log_n) +
           off;
  }
  __device__ const int64_t *key_tile(int pp, int i, int64_t off) const {
    const int kp = a.key_partition0 + pp;
    return a.key +
           ((int64_t(2 * kp) * a.key_chain + a.key_offset + i) << log_n) + off;
  }
  __device__ int64_t *acc(int i, int64_t off) const {
    return a.acc + ((int64_t(ct) * 2 * a.ext_limbs + i) << log_n) + off;
  }
};

// The batched key inner product with the small primes' products summed
// unreduced (ks_item_tile with WIDE) over `blocks` work items, in the
// block order of the batched kernel.
void launch_ks_inner_product_wide(const KsArgs &a, const DeviceTables &t,
                                  int64_t blocks, hipStream_t stream);

// Where one (ciphertext, partition)'s digits, constants and extended limbs
// live.
struct ExtItem {
  const int64_t *digits; // this partition's [maxz][N] Garner digits
  const int64_t *cst;    // this partition's ModUp constants
  int64_t *ext;          // this partition's [ext_limbs][N]
  int size, ext_limbs, moduli_id, ext_shoup_offset;
};

// OPAQUE: as for ks_item_tile.
template <int A, int MAXZ, bool SIGNED, bool OPAQUE = false>
__device__ __forceinline__ void ext_item_tile(const ExtItem &a,
                                              const DeviceTables &t,
                                              uint32_t *lds, int tile, int i) {
  const int64_t n = int64_t(1) << t.log_n;
  const int p = a.moduli_id + i;
  const int tid = OPAQUE ? opaque_tid() : int(threadIdx.x);
  const int c = tid & 15, r = tid >> 4;
  const int64_t sp = int64_t(tile) * 16 + c;
  const int size = a.size;
  const int64_t *dg = a.digits + sp;
  R16Pass s;
  s.tw = t.root + int64_t(p) * n;
  s.ts = t.root_shoup + int64_t(p) * n * 2;
  s.n = n;
  s.hi = 0;
  s.log_n = t.log_n;
  s.b = 8;
  s.r = r;
  s.q = t.q[p];
  s.k = t.k[p];
  s.q2 = int64_t(t.q2[p]);
  s.nq2 = A == kExact ? opaque_neg(s.q2) : 0;  // exact guards only
  s.nq = A == kExact ? 0 : uint64_t(opaque_neg(int64_t(s.q)));
  int64_t v[16];
#pragma unroll
  for (int m = 0; m < 16; m++) {
    int64_t d[MAXZ];
#pragma unroll
    for (int j = 0; j < MAXZ; j++)
      d[j] = j < size ? dg[int64_t(j) * n + (int64_t(r + 16 * m) << 8)] : 0;
    if (A == kExact)
      v[m] = extend_limb<MAXZ, SIGNED>(d, size, i, a.ext_limbs,
                                       a.cst + CkksParams::kKsExtOffset, s.q,
                                       s.k, s.q2);
    else
      v[m] = extend_limb_lazy<MAXZ>(
          d, size, i, a.ext_limbs,
          reinterpret_cast<const uint64_t *>(a.cst + a.ext_shoup_offset), s.q,
          s.nq, A == kLazySmall);
  }
  r16_forward_high<A>(v, s);
  r16_exchange_words(v, lds, c, r, true);
  r16_forward_low<A>(v, s);
  int64_t *out = a.ext + int64_t(i) * n + sp;
#pragma unroll
  for (int m = 0; m < 16; m++)
    out[int64_t(16 * r + m) << 8] = v[m];
}

} // namespace fhe
