// Kernels of the butterfly stage with a BootstrapKey (CKKSEngine::
// apply_butterfly with a PlainMatrix, src/ckks/engine_bootstrap.cpp:218-289),
// which works in the extended basis: the stage's chain limbs followed by the
// special primes, NTT form.
//
//   bk_rotate_kernel: ntt_rotate (src/ckks/core.cpp:964-997) of up to
//     kBkMaxTerms terms at once.  A term is the key inner product of a
//     decomposition with a fixed rotation key (CKKSCore::create_key_switcher,
//     src/ckks/core_key_switch.cpp:510-545), plus the source's polynomial 0
//     in half 0, permuted by the NTT-domain automorphism; or, for the
//     identity, the source itself.  The terms are written out one by one
//     (the baby steps) or summed (the giant steps of one stage).
//   bk_diagonal_kernel: the products of the baby-step texts with the stage's
//     encoded diagonals (multiply_for_butterfly, CKKSCore::multiply(ct, pt),
//     src/ckks/core.cpp:257-282) summed over each giant-step group
//     (add_inplace).
//   bk_attach_kernel: the chain limbs copied into the extended layout with
//     zero special limbs (attach_special_prime_chains_,
//     src/ckks/core.cpp:1127-1155, before its scalar product and NTT).
//
// Every value in the extended basis reaches an output only through ModDown,
// after an INTT whose result is reduced to [0, q) (CKKSCore::mod_down,
// src/ckks/core_key_switch.cpp:110-118): only residues matter there, so the
// kernels keep representatives below 2q (sums with one conditional
// subtraction of 2q, Montgomery products of an operand below 2q and one
// brought below q, each below 2q for q < 2^61).
//
// The NTT-domain automorphism maps every aligned block of 2^k slots onto an
// aligned block of 2^k slots (slot i holds the value at psi^(2 bitrev(i) +
// 1)): a wave's 64 gathered reads fall in one 512-byte block.
#include "kernels.h"

namespace fhe {

namespace {

// The source slot of output slot j under X -> X^g in NTT form.
__device__ __forceinline__ uint32_t ntt_source(uint32_t j, uint32_t g,
                                               int log_n) {
  const int rs = 32 - log_n;
  const uint32_t n = 1u << log_n;
  const uint32_t ei = 2 * (__builtin_bitreverse32(j) >> rs) + 1;
  const uint32_t ej = (g * ei) & (2 * n - 1);
  return __builtin_bitreverse32(ej >> 1) >> rs;
}

// a + b for a, b in [0, 2q): the result in [0, 2q)
__device__ __forceinline__ int64_t add_2q(int64_t a, int64_t b, int64_t q2) {
  const int64_t s = a + b;
  return s >= q2 ? s - q2 : s;
}

// Block w of the logical order, with consecutive logical blocks on one XCD
// (blocks are dealt to the 8 XCDs round-robin) when the grid is a multiple
// of 8: the blocks that read one key or diagonal tile share that XCD's L2.
__device__ __forceinline__ int64_t logical_block() {
  const int64_t grid = gridDim.x;
  if (grid % 8 != 0)
    return blockIdx.x;
  return int64_t(blockIdx.x & 7) * (grid >> 3) + (blockIdx.x >> 3);
}

template <bool SUM>
__global__ __launch_bounds__(kBlock) void bk_rotate_kernel(BkRotateArgs a,
                                                           DeviceTables t) {
  const int64_t n = int64_t(1) << t.log_n;
  const int64_t tiles = n / kBlock;
  // logical order (limb, tile, ciphertext), the ciphertext fastest
  const int64_t w = logical_block();
  const int c = int(w % a.count);
  const int64_t rest = w / a.count;
  const uint32_t j = uint32_t((rest % tiles) * kBlock + threadIdx.x);
  const int l = int(rest / tiles);
  const int p = a.moduli_id + l;
  const uint64_t q = t.q[p], k = t.k[p];
  const int64_t q2 = int64_t(t.q2[p]);
  const int64_t limb = int64_t(l) * n;
  const int64_t half = int64_t(a.ext_limbs) * n;
  int64_t s0 = 0, s1 = 0;
  if (SUM && a.accumulate) {
    const int64_t *o = a.out + int64_t(c) * a.out_c_stride + limb + j;
    s0 = o[0];
    s1 = o[half];
  }
  for (int x = 0; x < a.terms; x++) {
    const BkTerm &tm = a.term[x];
    const int64_t *src = tm.c0 + int64_t(c) * tm.src_stride + limb;
    int64_t y0, y1;
    if (tm.dec == nullptr) {
      y0 = src[j];
      y1 = tm.c1[int64_t(c) * tm.src_stride + limb + j];
    } else {
      const uint32_t s = ntt_source(j, tm.g, t.log_n);
      y0 = src[s];
      y1 = 0;
      const int64_t *d = tm.dec + int64_t(c) * tm.dec_stride + limb + s;
      const int64_t *kk = tm.key + limb + s;
      const int64_t ks = int64_t(tm.key_chain) * n;
      for (int pp = 0; pp < a.parts; pp++) {
        const int64_t dv = d[int64_t(pp) * half];
        // one operand below q keeps each product below 2q for q < 2^61
        const int64_t k0 = guard_u(kk[(2 * pp) * ks], int64_t(q));
        const int64_t k1 = guard_u(kk[(2 * pp + 1) * ks], int64_t(q));
        y0 = add_2q(y0, mont_mult_u(dv, k0, q, k), q2);
        y1 = add_2q(y1, mont_mult_u(dv, k1, q, k), q2);
      }
    }
    if (SUM) {
      s0 = add_2q(s0, y0, q2);
      s1 = add_2q(s1, y1, q2);
    } else {
      int64_t *o = a.out + int64_t(c) * a.out_c_stride +
                   int64_t(x) * a.out_t_stride + limb + j;
      o[0] = y0;
      o[half] = y1;
    }
  }
  if (SUM) {
    int64_t *o = a.out + int64_t(c) * a.out_c_stride + limb + j;
    o[0] = s0;
    o[half] = s1;
  }
}

__global__ __launch_bounds__(kBlock) void bk_diagonal_kernel(BkDiagArgs a,
                                                             DeviceTables t) {
  const int64_t n = int64_t(1) << t.log_n;
  const int64_t tiles = n / kBlock;
  const int64_t w = logical_block();
  const int c = int(w % a.count);
  const int64_t rest = w / a.count;
  const int64_t j = (rest % tiles) * kBlock + threadIdx.x;
  const int l = int(rest / tiles);
  const int p = a.moduli_id + l;
  const uint64_t q = t.q[p], k = t.k[p];
  const int64_t q2 = int64_t(t.q2[p]);
  const int64_t half = int64_t(a.ext_limbs) * n;
  const int64_t at = int64_t(l) * n + j;
  int64_t acc0[kBkMaxGroups], acc1[kBkMaxGroups];
#pragma unroll
  for (int g = 0; g < kBkMaxGroups; g++)
    acc0[g] = acc1[g] = 0;
  for (int i = 0; i < a.inner; i++) {
    const int64_t *src =
        i == 0 ? a.base + int64_t(c) * 2 * half
               : a.rotated + (int64_t(c) * (a.inner - 1) + (i - 1)) * 2 * half;
    const int64_t r0 = src[at], r1 = src[half + at];
#pragma unroll
    for (int g = 0; g < kBkMaxGroups; g++) {
      if (g < a.groups_here) {
        for (int m = a.member_start[g]; m < a.member_start[g + 1]; m++) {
          const int o = a.member[m];
          if (o == 0 && i == 0)
            continue;
          const int64_t dv =
              guard_u(a.diagonal[o * a.inner + i][at], int64_t(q));
          acc0[g] = add_2q(acc0[g], mont_mult_u(r0, dv, q, k), q2);
          acc1[g] = add_2q(acc1[g], mont_mult_u(r1, dv, q, k), q2);
        }
      }
    }
  }
  // sums [2][groups][count][E][N]
  const int64_t slot = int64_t(a.groups) * a.count * half;
#pragma unroll
  for (int g = 0; g < kBkMaxGroups; g++)
    if (g < a.groups_here) {
      int64_t *o = a.sums + (int64_t(a.group0 + g) * a.count + c) * half + at;
      o[0] = acc0[g];
      o[slot] = acc1[g];
    }
}

__global__ __launch_bounds__(kBlock) void bk_attach_kernel(
    const int64_t *in, int64_t *out, int64_t polys, int limbs, int ext_limbs,
    DeviceTables t) {
  const int64_t n = int64_t(1) << t.log_n;
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= (polys * ext_limbs) << t.log_n)
    return;
  const int64_t flat = idx >> t.log_n;
  const int64_t poly = flat / ext_limbs;
  const int l = int(flat - poly * ext_limbs);
  const int64_t j = idx & (n - 1);
  out[idx] = l < limbs ? in[(poly * limbs + l) * n + j] : 0;
}

} // namespace

bool bk_supported(const DeviceTables &t) { return t.log_n >= 8; }

void launch_bk_rotate(const DeviceTables &t, const BkRotateArgs &a, bool sum,
                      hipStream_t stream) {
  const int64_t blocks =
      int64_t(a.count) * a.ext_limbs * ((int64_t(1) << t.log_n) / kBlock);
  if (sum)
    hipLaunchKernelGGL(bk_rotate_kernel<true>, dim3(unsigned(blocks)),
                       dim3(kBlock), 0, stream, a, t);
  else
    hipLaunchKernelGGL(bk_rotate_kernel<false>, dim3(unsigned(blocks)),
                       dim3(kBlock), 0, stream, a, t);
}

void launch_bk_diagonal(const DeviceTables &t, const BkDiagArgs &a,
                        hipStream_t stream) {
  const int64_t blocks =
      int64_t(a.count) * a.ext_limbs * ((int64_t(1) << t.log_n) / kBlock);
  hipLaunchKernelGGL(bk_diagonal_kernel, dim3(unsigned(blocks)), dim3(kBlock),
                     0, stream, a, t);
}

void launch_bk_attach(const DeviceTables &t, const int64_t *in, int64_t *out,
                      int64_t polys, int limbs, int ext_limbs,
                      hipStream_t stream) {
  const int64_t total = (polys * ext_limbs) << t.log_n;
  hipLaunchKernelGGL(bk_attach_kernel, dim3(grid_for(total)), dim3(kBlock), 0,
                     stream, in, out, polys, limbs, ext_limbs, t);
}

} // namespace fhe
