// Work lists for batches of ciphertexts at uneven levels (N = 2^16).
//
// A batch of ciphertexts at different moduli_ids is processed one pipeline
// stage per launch: every kernel of the stage walks an explicit list of
// work units, each naming (ciphertext, polynomial or partition, limb,
// tile), so one launch covers all ciphertexts whatever their limb counts.
// The lists are built once per batch shape (the ciphertexts' moduli_ids)
// and kept on the device; per-call pointers travel as kernel arguments, so
// a launch sequence can be captured in a HIP graph and replayed.
//
// The kernels are persistent: a launch of G blocks walks its list with a
// stride of G (G below the unit count), and with `xcd` the list is cut into
// eight contiguous ranges, one per XCD (blocks are dealt round-robin to the
// eight XCDs), so that neighbouring units - the same prime's twiddles, the
// same key tile - meet in one L2.
#pragma once

#include "device.h"

namespace fhe {

constexpr int kUnevenMax = 64;      // ciphertexts per launch
constexpr int kUnevenMaxParts = 16; // key-switching partitions per ciphertext

// Static description of one ciphertext of a batch (device table).
struct UnevenCt {
  int moduli_id, limbs, ext, parts;
  int key_partition0; // key partition index of partition 0
  uint32_t signed_mask; // bit p: partition p takes the signed Garner forms
  int start[kUnevenMaxParts], size[kUnevenMaxParts];
  const int64_t *cst; // ModUp constants of (moduli_id, key moduli_id)
  int cst_stride, ext_shoup_offset;
};

// Per-ciphertext device pointers of one launch (a kernel argument).
struct UnevenPtrs {
  const int64_t *p[kUnevenMax];
};

// A work unit: four 8-bit fields.
__host__ __device__ inline uint32_t unit_pack(int ct, int a, int limb,
                                              int tile) {
  return uint32_t(ct) | (uint32_t(a) << 8) | (uint32_t(limb) << 16) |
         (uint32_t(tile) << 24);
}
__host__ __device__ inline int unit_ct(uint32_t u) { return int(u & 255); }
__host__ __device__ inline int unit_a(uint32_t u) { return int((u >> 8) & 255); }
__host__ __device__ inline int unit_limb(uint32_t u) {
  return int((u >> 16) & 255);
}
__host__ __device__ inline int unit_tile(uint32_t u) { return int(u >> 24); }

// One launch's slice of a work list and its schedule.
struct UnitList {
  const uint32_t *units;
  int count;
  int xcd; // eight contiguous ranges, one per XCD
};

// Launch schedule of one list: grid = min(count, cap) blocks (cap > 0).
struct UnevenSchedule {
  int cap = 0;   // 0: one block per unit
  int xcd = 1;
};

// Stage launchers (uneven_kernels.hip).  Arithmetic variants: 0 exact,
// 1 lazy small-prime, 2 lazy large-prime forms (ntt_r16.h).
//
// One 8-bit NTT pass over the listed limbs: unit (ct, poly, limb, tile) is
// limb `limb` of polynomial `poly` at src.p[ct] + (poly * stride + limb) N,
// stride = the ciphertext's limbs (ext_stride = 0) or extended limbs (1);
// prime moduli_id + limb.  inverse / b / pre / post / last as
// r16_limb_pass.
void launch_uneven_ntt(const DeviceTables &t, const UnevenCt *cts,
                       const UnitList &list, const UnevenPtrs &src,
                       const UnevenPtrs &dst, int ext_stride, bool inverse,
                       int b, const int64_t *pre, int post, bool last,
                       int variant, int cap, hipStream_t stream);

// Garner digits of every listed (ct, partition, coefficient chunk of 256):
// x.p[ct] is the INTT-form source polynomial, digits.p[ct] receives
// [parts][maxz][N].
void launch_uneven_digits(const DeviceTables &t, const UnevenCt *cts,
                          const UnitList &list, const UnevenPtrs &x,
                          const UnevenPtrs &digits, int maxz, int cap,
                          hipStream_t stream);

// Extended limbs (ct, partition, limb, tile) from the digits with the first
// forward NTT pass fused: ext.p[ct] receives [parts][ext][N].  variant 0 is
// the exact extension with signed digits.
void launch_uneven_extend(const DeviceTables &t, const UnevenCt *cts,
                          const UnitList &list, const UnevenPtrs &digits,
                          const UnevenPtrs &ext, int maxz, int variant,
                          int cap, hipStream_t stream);

// Key inner product (ct, 0, extended limb, tile) as ks_item_tile: ext.p[ct]
// [parts][ext][N] after the first pass, own.p[ct] the NTT-form source or
// null, acc.p[ct] [2][ext][N]; the key [key polys][key_chain][N] at
// key_moduli_id.
void launch_uneven_ks(const DeviceTables &t, const UnevenCt *cts,
                      const UnitList &list, const UnevenPtrs &ext,
                      const UnevenPtrs &own, const UnevenPtrs &acc,
                      const int64_t *key, int key_chain, int key_moduli_id,
                      bool lazy_accumulate, int variant, int cap,
                      hipStream_t stream);

// ModDown of both accumulators plus an addend (ct, half, 0, chunk of 256
// coefficients): out.p[ct][h] = guard(d_h + ModDown(acc_h)) with
// d_h = d.p[ct] + h * limbs * N for h < d_polys (else zero).
void launch_uneven_moddown(const DeviceTables &t, const UnevenCt *cts,
                           const UnitList &list, const UnevenPtrs &acc,
                           const UnevenPtrs &d, int d_polys,
                           const UnevenPtrs &out, int cap,
                           hipStream_t stream);

// Coefficient-domain automorphism X -> X^g of listed limbs (ct, poly, limb,
// chunk of 1024 coefficients) of duos: out.p[ct] = rotate(in.p[ct]).
void launch_uneven_automorphism(const DeviceTables &t, const UnevenCt *cts,
                                const UnitList &list, const UnevenPtrs &in,
                                const UnevenPtrs &out, uint32_t g, int cap,
                                hipStream_t stream);

} // namespace fhe
