// Fused key-switching core for N = 2^16 (two 8-bit NTT passes).
//
// The reference (src/ckks/core_key_switch.cpp:293-358, :110-160) runs, per
// partition, a full NTT of the extended polynomial, two Montgomery
// products with the key and an accumulate, and then a full INTT of both
// accumulators.  Here one workgroup owns a 4096-coefficient tile of one
// extended limb of one ciphertext and, for every partition in turn,
// finishes the forward NTT of that tile (the high-bit pass ran before),
// multiplies by both key halves and accumulates in registers; it then runs
// the first (low-bit) pass of the inverse NTT on both accumulators before a
// single write.  Each coefficient sees the reference's exact sequence of
// butterflies, products and guards.
#include "kernels.h"
#include "ks_tile.h"

namespace fhe {

namespace {


// One (extended limb, tile, ciphertext) work item with arithmetic A for the
// two transforms (the key products and the accumulation are the same in
// every form: Montgomery products of operands whose product stays below
// q * 2^62, accumulated with the reference's guards).
template <int A, bool PF>
__device__ __forceinline__ void ks_tile(const KsArgs &a, const DeviceTables &t,
                                        int64_t *lds, int ct, int tile,
                                        int i) {
  const int64_t n = int64_t(1) << t.log_n;
  const int p = a.moduli_id + i;
  const int tid = threadIdx.x;
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

  // block-uniform: every thread of the block has the same (i, pp)
  auto is_own = [&](int pp) {
    return a.own != nullptr && i >= a.start[pp] && i < a.start[pp] + a.size[pp];
  };
  auto ext_tile = [&](int pp) {
    return a.ext + ((int64_t(ct) * a.parts + pp) * a.ext_limbs + i) * n + off;
  };
  int64_t acc0[16], acc1[16], v[16];
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
      const int64_t *src =
          a.own + int64_t(ct) * a.own_stride + int64_t(i) * n + off + 16 * r;
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
    const int kp = a.key_partition0 + pp;
    const int64_t *k0 =
        a.key + (int64_t(2 * kp) * a.key_chain + a.key_offset + i) * n + off +
        16 * r;
    const int64_t *k1 = k0 + int64_t(a.key_chain) * n;
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

  if (A == kLazySmall && a.lazy_acc && a.parts > 1) {
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
  int64_t *d0 = a.acc + (int64_t(ct) * 2 * a.ext_limbs + i) * n + off;
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

template <bool PF>
__device__ __forceinline__ void ks_block(const KsArgs &a, const DeviceTables &t) {
  __shared__ int64_t lds[kR16Lds];
  const int64_t n = int64_t(1) << t.log_n;
  const int tiles = int(n >> 12);

  int64_t w = blockIdx.x;
  if (a.swizzle) {
    const int64_t per = int64_t(gridDim.x) >> 3;
    w = int64_t(blockIdx.x & 7) * per + (blockIdx.x >> 3);
  }
  const int ct = int(w % a.count);
  const int64_t rest = w / a.count;
  const int tile = int(rest % tiles);
  const int i = int(rest / tiles);
  const uint64_t q = t.q[a.moduli_id + i];
  if (!a.lazy)
    ks_tile<kExact, PF>(a, t, lds, ct, tile, i);
  else if (q < kLazySmallBound)
    ks_tile<kLazySmall, PF>(a, t, lds, ct, tile, i);
  else
    ks_tile<kLazyLarge, PF>(a, t, lds, ct, tile, i);
}

// ---------------------------------------------------------------------------
// ModUp fused with the first (high-bit) pass of the forward NTT of every
// extended limb.  A block owns one 4096-coefficient tile (16 columns of the
// stride-256 pass) of one ciphertext and one partition: it loads the
// partition's limbs at those coefficients, keeps the Garner digits of its
// 16 coefficients per thread in registers, and for each extended limb forms
// the limb values, runs the 8 high stages in registers and writes the tile
// once.  The extended polynomials never make the round trip through HBM
// between ModUp and the NTT.
struct ModupNttArgs {
  const int64_t *x;   // INTT-form source polynomial of ciphertext c at x + c*x_stride
  int64_t x_stride;
  int64_t *ext;       // [count][parts][ext_limbs][N]
  const int64_t *cst; // ModUp constants, partition p at cst + p * cst_stride
  int cst_stride;
  int count, parts, limbs, ext_limbs, moduli_id;
  int start[kMaxFusedParts]; // partition start limbs
  int size[kMaxFusedParts];
  uint32_t signed_mask; // bit p: partition p needs the signed products
};

// The digits of a thread's 16 coefficients, one 16-entry array per digit
// position (separate arrays stay in registers where one 2-D array of 64
// entries would be placed in scratch memory).
template <int MAXZ> struct DigitTile {
  int64_t d0[16], d1[16], d2[16], d3[16];
  __device__ __forceinline__ void set(int m, const int64_t (&v)[MAXZ]) {
    d0[m] = v[0];
    if (MAXZ > 1)
      d1[m] = v[MAXZ > 1 ? 1 : 0];
    if (MAXZ > 2) {
      d2[m] = v[MAXZ > 2 ? 2 : 0];
      d3[m] = v[MAXZ > 3 ? 3 : 0];
    }
  }
  __device__ __forceinline__ void get(int m, int64_t (&v)[MAXZ]) const {
    v[0] = d0[m];
    if (MAXZ > 1)
      v[MAXZ > 1 ? 1 : 0] = d1[m];
    if (MAXZ > 2) {
      v[MAXZ > 2 ? 2 : 0] = d2[m];
      v[MAXZ > 3 ? 3 : 0] = d3[m];
    }
  }
};

template <int A, int MAXZ, bool SIGNED>
__device__ __forceinline__ void modup_ntt_limb(const ModupNttArgs &a,
                                               const DeviceTables &t,
                                               int64_t *lds,
                                               const DigitTile<MAXZ> &d,
                                               int size, const int64_t *ec,
                                               int i, int64_t *dst, int c,
                                               int r, int64_t sp) {
  const int64_t n = int64_t(1) << t.log_n;
  const int p = a.moduli_id + i;
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
#pragma clang loop unroll(full)
  for (int m = 0; m < 16; m++) {
    int64_t dm[MAXZ];
    d.get(m, dm);
    v[m] = extend_limb<MAXZ, SIGNED>(dm, size, i, a.ext_limbs, ec, s.q, s.k,
                                     s.q2);
    if (MAXZ > 2)
      __builtin_amdgcn_sched_barrier(0);
  }
  r16_forward_high<A>(v, s);
  r16_exchange(v, lds, c, r, true);
  r16_forward_low<A>(v, s);
  int64_t *out = dst + int64_t(i) * n + sp;
#pragma unroll
  for (int m = 0; m < 16; m++)
    out[int64_t(16 * r + m) << 8] = v[m];
}

template <int MAXZ, bool SIGNED, bool LAZY>
__device__ __forceinline__ void modup_ntt_tile(const ModupNttArgs &a,
                                               const DeviceTables &t,
                                               int64_t *lds, int ct, int part,
                                               int tile) {
  const int64_t n = int64_t(1) << t.log_n;
  const int tid = threadIdx.x;
  const int c = tid & 15, r = tid >> 4;
  const int64_t sp = int64_t(tile) * 16 + c; // column of the stride-256 pass
  const int start = a.start[part], size = a.size[part];
  const int64_t *cst = a.cst + int64_t(part) * a.cst_stride;
  const int64_t *xs = a.x + int64_t(ct) * a.x_stride + sp;

  // Garner digits of this thread's 16 coefficients (r + 16 m) << 8 | sp
  DigitTile<MAXZ> d;
#pragma clang loop unroll(full)
  for (int m = 0; m < 16; m++) {
    int64_t xv[MAXZ], dm[MAXZ];
#pragma unroll
    for (int j = 0; j < MAXZ; j++)
      xv[j] = j < size ? xs[int64_t(start + j) * n + (int64_t(r + 16 * m) << 8)]
                       : 0;
    garner_digits<MAXZ, SIGNED>(xv, dm, size, a.moduli_id + start, cst, t);
    d.set(m, dm);
    // one coefficient's digit chain at a time: interleaving all sixteen
    // would need far more registers than the tile of digits itself
    __builtin_amdgcn_sched_barrier(0);
  }
  const int64_t *ec = cst + CkksParams::kKsExtOffset;
  int64_t *dst =
      a.ext + (int64_t(ct) * a.parts + part) * int64_t(a.ext_limbs) * n;
  for (int i = 0; i < a.ext_limbs; i++) {
    if (!LAZY)
      modup_ntt_limb<kExact, MAXZ, SIGNED>(a, t, lds, d, size, ec, i, dst, c,
                                           r, sp);
    else if (t.q[a.moduli_id + i] < kLazySmallBound)
      modup_ntt_limb<kLazySmall, MAXZ, SIGNED>(a, t, lds, d, size, ec, i, dst,
                                               c, r, sp);
    else
      modup_ntt_limb<kLazyLarge, MAXZ, SIGNED>(a, t, lds, d, size, ec, i, dst,
                                               c, r, sp);
  }
}

template <int MAXZ, bool LAZY>
__global__ __launch_bounds__(kBlock, 1) void modup_ntt_r16_kernel(
    ModupNttArgs a, DeviceTables t) {
  __shared__ int64_t lds[kR16Lds];
  const int tiles = int((int64_t(1) << t.log_n) >> 12);
  // block -> (ciphertext, partition, tile), tile fastest: consecutive blocks
  // read neighbouring columns of the same source limbs
  const int tile = int(blockIdx.x % tiles);
  const int64_t rest = blockIdx.x / tiles;
  const int part = int(rest % a.parts);
  const int ct = int(rest / a.parts);
  if ((a.signed_mask >> part) & 1)
    modup_ntt_tile<MAXZ, true, LAZY>(a, t, lds, ct, part, tile);
  else
    modup_ntt_tile<MAXZ, false, LAZY>(a, t, lds, ct, part, tile);
}

// ---------------------------------------------------------------------------
// ModUp in two stages around HBM: the Garner digits of every partition
// (MAXZ limbs per partition, far fewer than the extended limbs), then one
// kernel that forms each extended limb from the digits and runs the first
// (high-bit) forward NTT pass on it before its only write.  The 4096-
// coefficient digit tile of a block is re-read by the blocks of the other
// extended limbs; the limb index runs fastest in the block order so those
// reads hit the XCD's L2.
struct DigitArgs {
  const int64_t *x;   // source polynomial of ciphertext c at x + c * x_stride
  int64_t x_stride;
  int64_t *digits;    // [count][parts][MAXZ][N]
  const int64_t *cst;
  int cst_stride;
  int count, parts, moduli_id, maxz;
  int start[kMaxFusedParts];
  int size[kMaxFusedParts];
  uint32_t signed_mask;
};

template <int MAXZ>
__global__ __launch_bounds__(kBlock) void modup_digits_kernel(DigitArgs a,
                                                             DeviceTables t) {
  const int64_t n = int64_t(1) << t.log_n;
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= int64_t(a.count) * a.parts * n)
    return;
  const int64_t coef = idx & (n - 1);
  const int64_t cp = idx >> t.log_n; // ciphertext * parts + partition
  const int part = int(cp % a.parts);
  const int64_t ct = cp / a.parts;
  const int start = a.start[part], size = a.size[part];
  const int64_t *xs = a.x + ct * a.x_stride + coef;
  const int64_t *cst = a.cst + int64_t(part) * a.cst_stride;
  int64_t xv[MAXZ], d[MAXZ];
#pragma unroll
  for (int j = 0; j < MAXZ; j++)
    xv[j] = j < size ? xs[int64_t(start + j) * n] : 0;
  if ((a.signed_mask >> part) & 1)
    garner_digits<MAXZ, true>(xv, d, size, a.moduli_id + start, cst, t);
  else
    garner_digits<MAXZ, false>(xv, d, size, a.moduli_id + start, cst, t);
  int64_t *out = a.digits + cp * MAXZ * n + coef;
#pragma unroll
  for (int j = 0; j < MAXZ; j++)
    if (j < size)
      out[int64_t(j) * n] = d[j];
}

// modup_digits_kernel over kChunk coefficients of one (ciphertext,
// partition) per block (chunked_ok): four independent Garner chains per
// thread.  A partition's limbs past its size are loaded again as its last
// limb and unused, so that every load is unconditional.
template <int MAXZ>
__global__ __launch_bounds__(kBlock) void modup_digits_chunk_kernel(
    DigitArgs a, DeviceTables t) {
  const int64_t n = int64_t(1) << t.log_n;
  const int64_t e0 = int64_t(blockIdx.x) * kChunk;
  const uint32_t cp = uint32_t(e0 >> t.log_n); // ciphertext * parts + part
  const int part = int(cp % uint32_t(a.parts));
  const int64_t ct = cp / uint32_t(a.parts);
  const int64_t coef = (e0 & (n - 1)) + int64_t(threadIdx.x) * kPerThread;
  const int start = a.start[part], size = a.size[part];
  const int64_t *xs = a.x + ct * a.x_stride + coef;
  const int64_t *cst = a.cst + int64_t(part) * a.cst_stride;
  int64_t xv[kPerThread][MAXZ], d[kPerThread][MAXZ];
#pragma unroll
  for (int j = 0; j < MAXZ; j++) {
    int64_t v[kPerThread];
    load4(xs + int64_t(start + min(j, size - 1)) * n, v);
#pragma unroll
    for (int g = 0; g < kPerThread; g++)
      xv[g][j] = j < size ? v[g] : 0;
  }
  if ((a.signed_mask >> part) & 1) {
#pragma unroll
    for (int g = 0; g < kPerThread; g++)
      garner_digits<MAXZ, true>(xv[g], d[g], size, a.moduli_id + start, cst,
                                t);
  } else {
#pragma unroll
    for (int g = 0; g < kPerThread; g++)
      garner_digits<MAXZ, false>(xv[g], d[g], size, a.moduli_id + start, cst,
                                 t);
  }
  int64_t *out = a.digits + int64_t(cp) * MAXZ * n + coef;
#pragma unroll
  for (int j = 0; j < MAXZ; j++)
    if (j < size) {
      int64_t v[kPerThread];
#pragma unroll
      for (int g = 0; g < kPerThread; g++)
        v[g] = d[g][j];
      store4(out + int64_t(j) * n, v);
    }
}

struct ExtendArgs {
  const int64_t *digits; // [count][parts][maxz][N]
  int64_t *ext;          // [count][parts][ext_limbs][N]
  const int64_t *cst;
  int cst_stride, ext_shoup_offset;
  int count, parts, ext_limbs, moduli_id, maxz, swizzle;
  int size[kMaxFusedParts];
  int start[kMaxFusedParts];
  int skip_own;          // leave each partition's own limbs unwritten
  uint32_t signed_mask;
};

template <int A, int MAXZ, bool SIGNED>
__device__ __forceinline__ void extend_ntt_tile(const ExtendArgs &a,
                                                const DeviceTables &t,
                                                uint32_t *lds, int64_t cp,
                                                int part, int tile, int i) {
  const int64_t n = int64_t(1) << t.log_n;
  ExtItem it;
  it.digits = a.digits + cp * a.maxz * n;
  it.cst = a.cst + int64_t(part) * a.cst_stride;
  it.ext = a.ext + cp * a.ext_limbs * n;
  it.size = a.size[part];
  it.ext_limbs = a.ext_limbs;
  it.moduli_id = a.moduli_id;
  it.ext_shoup_offset = a.ext_shoup_offset;
  ext_item_tile<A, MAXZ, SIGNED>(it, t, lds, tile, i);
}

template <int MAXZ, bool LAZY>
__global__ __launch_bounds__(kBlock) void extend_ntt_r16_kernel(
    ExtendArgs a, DeviceTables t) {
  __shared__ uint32_t lds[kR16LdsWords];
  const int tiles = int((int64_t(1) << t.log_n) >> 12);
  int64_t w = blockIdx.x;
  if (a.swizzle) {
    const int64_t per = int64_t(gridDim.x) >> 3;
    w = int64_t(blockIdx.x & 7) * per + (blockIdx.x >> 3);
  }
  // extended limb fastest, then tile, then (ciphertext, partition)
  const int i = int(w % a.ext_limbs);
  const int64_t rest = w / a.ext_limbs;
  const int tile = int(rest % tiles);
  const int64_t cp = rest / tiles;
  const int part = int(cp % a.parts);
  // the whole block leaves (before any barrier) for a partition's own limb
  if (a.skip_own && i >= a.start[part] && i < a.start[part] + a.size[part])
    return;
  const bool sgn = (a.signed_mask >> part) & 1;
  if (!LAZY) {
    if (sgn)
      extend_ntt_tile<kExact, MAXZ, true>(a, t, lds, cp, part, tile, i);
    else
      extend_ntt_tile<kExact, MAXZ, false>(a, t, lds, cp, part, tile, i);
  } else if (t.q[a.moduli_id + i] < kLazySmallBound) {
    extend_ntt_tile<kLazySmall, MAXZ, false>(a, t, lds, cp, part, tile, i);
  } else {
    extend_ntt_tile<kLazyLarge, MAXZ, false>(a, t, lds, cp, part, tile, i);
  }
}

template <bool PF>
__global__ __launch_bounds__(kBlock) void ks_inner_product_r16_kernel(
    KsArgs a, DeviceTables t) {
  ks_block<PF>(a, t);
}

// The same with the register allocation held to three waves per SIMD (part
// of the loop's registers spill to scratch in exchange for occupancy).
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3)))
void ks_inner_product_r16_occ3_kernel(KsArgs a, DeviceTables t) {
  ks_block<false>(a, t);
}

} // namespace

bool ks_fused_supported(const DeviceTables &t) { return t.log_n == 16; }

bool staged_modup_supported(const DeviceTables &t, int max_partition_size,
                            int parts) {
  return t.log_n == 16 && max_partition_size <= 4 && parts <= kMaxFusedParts;
}

void launch_staged_modup(const DeviceTables &t, const int64_t *x,
                         int64_t x_stride, int64_t *digits, int64_t *ext,
                         int count, int limbs, int moduli_id,
                         const int *starts, const int *sizes,
                         const bool *is_signed, int parts,
                         const int64_t *constants, int constants_stride,
                         int ext_shoup_offset, bool lazy, bool skip_own,
                         hipStream_t stream) {
  int maxz = 1;
  uint32_t mask = 0;
  bool any_signed = false;
  for (int p = 0; p < parts; p++) {
    while (maxz < sizes[p])
      maxz <<= 1;
    mask |= is_signed[p] ? (uint32_t(1) << p) : 0u;
    any_signed = any_signed || is_signed[p];
  }
  DigitArgs d{};
  d.x = x;
  d.x_stride = x_stride;
  d.digits = digits;
  d.cst = constants;
  d.cst_stride = constants_stride;
  d.count = count;
  d.parts = parts;
  d.moduli_id = moduli_id;
  d.maxz = maxz;
  d.signed_mask = mask;
  ExtendArgs e{};
  e.digits = digits;
  e.ext = ext;
  e.cst = constants;
  e.cst_stride = constants_stride;
  e.ext_shoup_offset = ext_shoup_offset;
  e.count = count;
  e.parts = parts;
  e.ext_limbs = limbs + t.alpha;
  e.moduli_id = moduli_id;
  e.maxz = maxz;
  e.signed_mask = mask;
  e.skip_own = skip_own ? 1 : 0;
  for (int p = 0; p < parts; p++) {
    d.start[p] = starts[p];
    d.size[p] = sizes[p];
    e.size[p] = sizes[p];
    e.start[p] = starts[p];
  }
  // the lazy extension needs canonical digits: only for non-negative ones
  const bool use_lazy = lazy && !any_signed;
  const int tiles = int((int64_t(1) << t.log_n) >> 12);
  const int64_t eblocks = int64_t(count) * parts * tiles * e.ext_limbs;
  e.swizzle = eblocks % 8 == 0 ? 1 : 0;
  const dim3 dgrid(grid_for((int64_t(count) * parts) << t.log_n));
  const dim3 egrid{unsigned(eblocks), 1, 1};
  const bool chunked = chunked_ok(t.log_n, {x, digits}) && x_stride % 2 == 0;
#define FHE_STAGED(W)                                                          \
  do {                                                                         \
    if (chunked)                                                               \
      hipLaunchKernelGGL(modup_digits_chunk_kernel<W>,                         \
                         dim3(unsigned(((int64_t(count) * parts) << t.log_n) / \
                                       kChunk)),                               \
                         dim3(kBlock), 0, stream, d, t);                       \
    else                                                                       \
      hipLaunchKernelGGL(modup_digits_kernel<W>, dgrid, dim3(kBlock), 0,      \
                         stream, d, t);                                        \
    if (use_lazy)                                                              \
      hipLaunchKernelGGL((extend_ntt_r16_kernel<W, true>), egrid, dim3(kBlock), \
                         0, stream, e, t);                                     \
    else                                                                       \
      hipLaunchKernelGGL((extend_ntt_r16_kernel<W, false>), egrid,             \
                         dim3(kBlock), 0, stream, e, t);                       \
  } while (0)
  if (maxz == 1)
    FHE_STAGED(1);
  else if (maxz == 2)
    FHE_STAGED(2);
  else
    FHE_STAGED(4);
#undef FHE_STAGED
}

bool modup_ntt_fused_supported(const DeviceTables &t, int max_partition_size,
                               int parts) {
  return t.log_n == 16 && max_partition_size <= 4 && parts <= kMaxFusedParts;
}

void launch_modup_ntt_fused(const DeviceTables &t, const int64_t *x,
                            int64_t x_stride, int64_t *ext, int count,
                            int limbs, int moduli_id, const int *starts,
                            const int *sizes, const bool *is_signed, int parts,
                            const int64_t *constants, int constants_stride,
                            bool lazy, hipStream_t stream) {
  ModupNttArgs a{};
  a.x = x;
  a.x_stride = x_stride;
  a.ext = ext;
  a.cst = constants;
  a.cst_stride = constants_stride;
  a.count = count;
  a.parts = parts;
  a.limbs = limbs;
  a.ext_limbs = limbs + t.alpha;
  a.moduli_id = moduli_id;
  int max_size = 0;
  for (int p = 0; p < parts; p++) {
    a.start[p] = starts[p];
    a.size[p] = sizes[p];
    a.signed_mask |= is_signed[p] ? (uint32_t(1) << p) : 0u;
    max_size = sizes[p] > max_size ? sizes[p] : max_size;
  }
  const int tiles = int((int64_t(1) << t.log_n) >> 12);
  const dim3 grid(unsigned(int64_t(count) * parts * tiles));
#define FHE_MODUP_NTT(W, L)                                                    \
  hipLaunchKernelGGL((modup_ntt_r16_kernel<W, L>), grid, dim3(kBlock), 0,      \
                     stream, a, t)
  if (max_size <= 1) {
    if (lazy)
      FHE_MODUP_NTT(1, true);
    else
      FHE_MODUP_NTT(1, false);
  } else if (max_size <= 2) {
    if (lazy)
      FHE_MODUP_NTT(2, true);
    else
      FHE_MODUP_NTT(2, false);
  } else {
    if (lazy)
      FHE_MODUP_NTT(4, true);
    else
      FHE_MODUP_NTT(4, false);
  }
#undef FHE_MODUP_NTT
}

void launch_ks_inner_product_fused(const DeviceTables &t, const int64_t *ext,
                                   const int64_t *key, int key_chain,
                                   int key_offset, int key_partition0,
                                   int parts, int64_t *acc, int count,
                                   int ext_limbs, int moduli_id, bool lazy,
                                   const int64_t *own, int64_t own_stride,
                                   const int *starts, const int *sizes,
                                   bool lazy_accumulate, bool wide_accumulate,
                                   bool prefetch, int occupancy,
                                   hipStream_t stream) {
  const int tiles = int((int64_t(1) << t.log_n) >> 12);
  const int64_t blocks = int64_t(ext_limbs) * tiles * count;
  KsArgs a{};
  a.ext = ext;
  a.key = key;
  a.acc = acc;
  a.count = count;
  a.parts = parts;
  a.ext_limbs = ext_limbs;
  a.moduli_id = moduli_id;
  a.key_chain = key_chain;
  a.key_offset = key_offset;
  a.key_partition0 = key_partition0;
  a.swizzle = blocks % 8 == 0 ? 1 : 0;
  a.lazy = lazy ? 1 : 0;
  a.own = own;
  a.own_stride = own_stride;
  a.lazy_acc = lazy_accumulate ? 1 : 0;
  for (int p = 0; p < parts && p < kMaxFusedParts; p++) {
    a.start[p] = own ? starts[p] : 0;
    a.size[p] = own ? sizes[p] : 0;
  }
  // small primes' key products summed unreduced (a kernel of its own
  // translation unit, ks_wide.hip)
  if (wide_accumulate && lazy && !prefetch && occupancy < 3 && parts > 1 &&
      parts <= kMaxWideParts) {
    launch_ks_inner_product_wide(a, t, blocks, stream);
    return;
  }
  if (prefetch)
    hipLaunchKernelGGL(ks_inner_product_r16_kernel<true>,
                       dim3(unsigned(blocks)), dim3(kBlock), 0, stream, a, t);
  else if (occupancy >= 3)
    hipLaunchKernelGGL(ks_inner_product_r16_occ3_kernel,
                       dim3(unsigned(blocks)), dim3(kBlock), 0, stream, a, t);
  else
    hipLaunchKernelGGL(ks_inner_product_r16_kernel<false>,
                       dim3(unsigned(blocks)), dim3(kBlock), 0, stream, a, t);
}

} // namespace fhe
