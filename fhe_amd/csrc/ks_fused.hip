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
#include "ntt_r16.h"

namespace fhe {

namespace {

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
};

// One (extended limb, tile, ciphertext) work item with arithmetic A for the
// two transforms (the key products and the accumulation are the same in
// every form: Montgomery products of operands whose product stays below
// q * 2^62, accumulated with the reference's guards).
template <int A>
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
  const int64_t qq = int64_t(f.q);

  int64_t acc0[16], acc1[16], v[16];
  for (int pp = 0; pp < a.parts; pp++) {
    const int64_t *src =
        a.ext + ((int64_t(ct) * a.parts + pp) * a.ext_limbs + i) * n + off;
#pragma unroll
    for (int m = 0; m < 16; m++)
      v[m] = src[r + 16 * m];
    r16_forward_high<A>(v, f);
    r16_exchange(v, lds, c, r, true);
    r16_forward_low<A>(v, f);
    r16_forward_settle<A>(v, f);
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

  R16Pass g = f;
  g.tw = t.root_inverse + int64_t(p) * n;
  g.ts = t.root_inverse_shoup + int64_t(p) * n * 2;
  int64_t *d0 = a.acc + (int64_t(ct) * 2 * a.ext_limbs + i) * n + off;
  int64_t *d1 = d0 + int64_t(a.ext_limbs) * n;
  r16_inverse_low<A>(acc0, g);
  r16_exchange(acc0, lds, c, r, false);
  r16_inverse_high<A>(acc0, g);
#pragma unroll
  for (int m = 0; m < 16; m++)
    d0[r + 16 * m] = acc0[m];
  r16_inverse_low<A>(acc1, g);
  r16_exchange(acc1, lds, c, r, false);
  r16_inverse_high<A>(acc1, g);
#pragma unroll
  for (int m = 0; m < 16; m++)
    d1[r + 16 * m] = acc1[m];
}

__global__ __launch_bounds__(kBlock) void ks_inner_product_r16_kernel(
    KsArgs a, DeviceTables t) {
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
    ks_tile<kExact>(a, t, lds, ct, tile, i);
  else if (q < kLazySmallBound)
    ks_tile<kLazySmall>(a, t, lds, ct, tile, i);
  else
    ks_tile<kLazyLarge>(a, t, lds, ct, tile, i);
}

} // namespace

bool ks_fused_supported(const DeviceTables &t) { return t.log_n == 16; }

void launch_ks_inner_product_fused(const DeviceTables &t, const int64_t *ext,
                                   const int64_t *key, int key_chain,
                                   int key_offset, int key_partition0,
                                   int parts, int64_t *acc, int count,
                                   int ext_limbs, int moduli_id, bool lazy,
                                   hipStream_t stream) {
  const int tiles = int((int64_t(1) << t.log_n) >> 12);
  const int64_t blocks = int64_t(ext_limbs) * tiles * count;
  KsArgs a{ext,        key,       acc,        count,
           parts,      ext_limbs, moduli_id,  key_chain,
           key_offset, key_partition0, blocks % 8 == 0 ? 1 : 0,
           lazy ? 1 : 0};
  hipLaunchKernelGGL(ks_inner_product_r16_kernel, dim3(unsigned(blocks)),
                     dim3(kBlock), 0, stream, a, t);
}

} // namespace fhe
