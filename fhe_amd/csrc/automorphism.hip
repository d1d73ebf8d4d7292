// Automorphisms X -> X^g of the negacyclic ring, g odd below 2N.
//
// Coefficient domain (the reference's rotate_, src/ckks/core.cpp:845-869,
// with the permutation and sign of src/ckks/encode.cpp:513-526): input
// coefficient i lands at (g i mod 2N) mod N, negated when g i mod 2N >= N,
// and every output passes make_unsigned_reduce_2q_to_q
// (src/montgomery/compound.cpp:85-91).  Here each thread owns one OUTPUT
// coefficient t and gathers its source: with h = g^-1 mod 2N and
// s = h t mod 2N, the source is s (sign +) when s < N and s - N (sign -)
// otherwise, because g N = N mod 2N for odd g.
//
// NTT domain (src/ckks/core.cpp:886-910): NTT slot i holds the value at
// psi^(2 bitrev(i) + 1), so the automorphism is the pure gather
// out[i] = in[bitrev((g (2 bitrev(i) + 1) mod 2N - 1) / 2)].
//
// Both kernels read one limb at random positions and write it in order.
// Blocks are dealt so that the blocks of a limb run on one XCD whenever the
// grid allows, which keeps the gathered limb (N * 8 bytes) in that XCD's L2;
// a block's 1024 outputs lie in one limb when N >= 1024.
#include "kernels.h"

namespace fhe {

namespace {

constexpr int kPerThread = 4;
constexpr int kTile = kBlock * kPerThread;

// Block w of the logical order runs on XCD w / (grid / 8).
__device__ __forceinline__ int64_t xcd_block(int swizzle) {
  if (!swizzle)
    return blockIdx.x;
  const int64_t per = int64_t(gridDim.x) >> 3;
  return int64_t(blockIdx.x & 7) * per + (blockIdx.x >> 3);
}

// out never overlaps in (a permutation cannot run in place); with N a
// multiple of kTile a block's coefficients lie in one limb, whose prime is
// then read once and the block's gathers issue together.
__global__ __launch_bounds__(kBlock) void automorphism_kernel(
    const int64_t *__restrict__ in, int64_t *__restrict__ out, int64_t total,
    int limbs, int moduli_id, uint32_t h, int swizzle, DeviceTables t) {
  const uint32_t n = 1u << t.log_n;
  const int64_t base = xcd_block(swizzle) * kTile + threadIdx.x;
  if (n % kTile == 0 && base < total) {
    const int64_t f = base >> t.log_n; // flat limb, one per block
    const int64_t q =
        int64_t(t.q[moduli_id + int(uint32_t(f) % uint32_t(limbs))]);
    int64_t x[kPerThread];
    uint32_t s[kPerThread];
#pragma unroll
    for (int e = 0; e < kPerThread; e++) {
      const uint32_t o = uint32_t(base + e * kBlock) & (n - 1);
      // mod 2N of a product that wrapped mod 2^32 is still exact
      s[e] = (h * o) & (2 * n - 1);
      x[e] = in[f * n + (s[e] & (n - 1))];
    }
#pragma unroll
    for (int e = 0; e < kPerThread; e++) {
      const int64_t y = (s[e] & n) ? -x[e] : x[e];
      out[base + e * kBlock] = y < 0 ? y + q : guard(y, q);
    }
    return;
  }
#pragma unroll
  for (int e = 0; e < kPerThread; e++) {
    const int64_t idx = base + e * kBlock;
    if (idx >= total)
      return;
    const int64_t f = idx >> t.log_n; // flat limb
    const uint32_t o = uint32_t(idx) & (n - 1);
    const int64_t q = int64_t(t.q[moduli_id + int(f % limbs)]);
    const uint32_t s = (h * o) & (2 * n - 1);
    const int64_t x = in[f * n + (s & (n - 1))];
    const int64_t y = (s & n) ? -x : x;
    out[idx] = y < 0 ? y + q : guard(y, q);
  }
}

__global__ __launch_bounds__(kBlock) void ntt_automorphism_kernel(
    const int64_t *in, int64_t *out, int64_t total, uint32_t g, int swizzle,
    DeviceTables t) {
  const uint32_t n = 1u << t.log_n;
  const int rs = 32 - t.log_n;
  const int64_t base = xcd_block(swizzle) * kTile + threadIdx.x;
#pragma unroll
  for (int e = 0; e < kPerThread; e++) {
    const int64_t idx = base + e * kBlock;
    if (idx >= total)
      return;
    const int64_t f = idx >> t.log_n;
    const uint32_t o = uint32_t(idx) & (n - 1);
    const uint32_t ei = 2 * (__builtin_bitreverse32(o) >> rs) + 1;
    const uint32_t ej = (g * ei) & (2 * n - 1);
    out[idx] = in[f * n + (__builtin_bitreverse32(ej >> 1) >> rs)];
  }
}

// g^-1 mod 2^32 for odd g (Newton: each step doubles the correct bits).
uint32_t odd_inverse(uint32_t g) {
  uint32_t x = g; // correct to 3 bits
  for (int i = 0; i < 4; i++)
    x *= 2u - g * x;
  return x;
}

dim3 limb_grid(const DeviceTables &t, int64_t total, int &swizzle) {
  const int64_t blocks = (total + kTile - 1) / kTile;
  swizzle = blocks % 8 == 0 ? 1 : 0;
  return dim3(unsigned(blocks));
}

} // namespace

void launch_automorphism(const DeviceTables &t, const int64_t *in,
                         int64_t *out, int64_t total_limbs, int limbs,
                         int moduli_id, uint32_t g, hipStream_t stream) {
  const uint32_t two_n = 2u << t.log_n;
  const uint32_t h = odd_inverse(g) & (two_n - 1);
  const int64_t total = total_limbs << t.log_n;
  int swizzle = 0;
  const dim3 grid = limb_grid(t, total, swizzle);
  hipLaunchKernelGGL(automorphism_kernel, grid, dim3(kBlock), 0, stream, in,
                     out, total, limbs, moduli_id, h, swizzle, t);
}

void launch_ntt_automorphism(const DeviceTables &t, const int64_t *in,
                             int64_t *out, int64_t total_limbs, uint32_t g,
                             hipStream_t stream) {
  const int64_t total = total_limbs << t.log_n;
  int swizzle = 0;
  const dim3 grid = limb_grid(t, total, swizzle);
  hipLaunchKernelGGL(ntt_automorphism_kernel, grid, dim3(kBlock), 0, stream,
                     in, out, total, g, swizzle, t);
}

} // namespace fhe
