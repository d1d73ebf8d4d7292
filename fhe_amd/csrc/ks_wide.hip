// The batched key inner product (N = 2^16) with the key products of the
// small primes summed as unreduced 96-bit values and reduced once per
// accumulator (ks_item_tile with WIDE, ks_tile.h).  It sits in a translation
// unit of its own: next to the other instantiations of the loop in
// ks_fused.hip the compiler allocated the batched kernel more registers
// (one wave per SIMD instead of two).
#include "ks_tile.h"

namespace fhe {

namespace {

__global__ __launch_bounds__(kBlock) void ks_inner_product_wide_kernel(
    KsArgs a, DeviceTables t) {
  __shared__ int64_t lds[kR16Lds];
  const int64_t n = int64_t(1) << t.log_n;
  const int tiles = int(n >> 12);
  int64_t w = blockIdx.x;
  if (a.swizzle) {
    const int64_t per = int64_t(gridDim.x) >> 3;
    w = int64_t(blockIdx.x & 7) * per + (blockIdx.x >> 3);
  }
  // block -> (limb i, tile, ciphertext), ciphertext fastest: the blocks of a
  // key tile run together and share its reads through L2
  const int ct = int(w % a.count);
  const int64_t rest = w / a.count;
  const int tile = int(rest % tiles);
  const int i = int(rest / tiles);
  const KsBatchAddr ad(a, ct, t.log_n);
  if (t.q[a.moduli_id + i] < kLazySmallBound)
    ks_item_tile<kLazySmall, false, true>(ad, t, lds, tile, i);
  else
    ks_item_tile<kLazyLarge, false, false>(ad, t, lds, tile, i);
}

} // namespace

void launch_ks_inner_product_wide(const KsArgs &a, const DeviceTables &t,
                                  int64_t blocks, hipStream_t stream) {
  hipLaunchKernelGGL(ks_inner_product_wide_kernel, dim3(unsigned(blocks)),
                     dim3(kBlock), 0, stream, a, t);
}

} // namespace fhe
