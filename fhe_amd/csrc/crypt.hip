// Device arithmetic of public-key encryption (CKKSCore::encrypt(pt, pk),
// src/ckks/core.cpp:35-82).  The samples (binary u, errors e0, e1) are drawn
// on the host, one N-vector of small signed integers each per ciphertext;
// these kernels lift them to every limb (tile_unsigned_reduce_2q_to_q_,
// src/ckks/core_calculate.cpp:485-531) and fold the errors and the
// plaintext into the key products in one pass.
#include "kernels.h"

namespace fhe {

namespace {

__device__ __forceinline__ int64_t unsigned_reduce(int64_t x, int64_t q) {
  return x < 0 ? x + q : guard(x, q);
}

// out[c][l][i] = make_unsigned_reduce(s[c][i], q_l)
__global__ __launch_bounds__(kBlock) void tile_unsigned_kernel(
    const int64_t *s, int64_t *out, int64_t total, int limbs, int moduli_id,
    DeviceTables t) {
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= total)
    return;
  const int64_t n = int64_t(1) << t.log_n;
  const int64_t flat = idx >> t.log_n;
  const int64_t c = flat / limbs;
  const int l = int(flat - c * limbs);
  out[idx] = unsigned_reduce(s[c * n + (idx & (n - 1))],
                             int64_t(t.q[moduli_id + l]));
}

// out[l][i] = make_unsigned_reduce(value, q_l) at i = pos, 0 elsewhere: the
// tiled sample of a monomial value * X^pos, written on the device
__global__ __launch_bounds__(kBlock) void tile_monomial_kernel(
    int64_t *out, int64_t total, int moduli_id, int64_t value, int64_t pos,
    DeviceTables t) {
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= total)
    return;
  const int l = int(idx >> t.log_n);
  const int64_t i = idx & ((int64_t(1) << t.log_n) - 1);
  out[idx] = i == pos ? unsigned_reduce(value, int64_t(t.q[moduli_id + l]))
                      : 0;
}

// c0 = reduce(c0 + reduce(pt + tile(e0))), c1 = reduce(c1 + tile(e1)) with
// reduce(x) = guard(guard(x, 2q), q) (mont_add_reduce_2q_to_q).
__global__ __launch_bounds__(kBlock) void encrypt_finish_kernel(
    const int64_t *pt, const int64_t *e0, const int64_t *e1, int64_t *out,
    int64_t total, int limbs, int moduli_id, DeviceTables t) {
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= total)
    return;
  const int64_t n = int64_t(1) << t.log_n;
  const int64_t poly = int64_t(limbs) * n;
  const int64_t c = idx / poly;
  const int64_t r = idx - c * poly;
  const int p = moduli_id + int(r >> t.log_n);
  const int64_t q = int64_t(t.q[p]), q2 = int64_t(t.q2[p]);
  const int64_t i = r & (n - 1);
  const int64_t err0 =
      guard(guard(pt[idx] + unsigned_reduce(e0[c * n + i], q), q2), q);
  const int64_t err1 = unsigned_reduce(e1[c * n + i], q);
  int64_t *o = out + c * 2 * poly + r;
  o[0] = guard(guard(o[0] + err0, q2), q);
  o[poly] = guard(guard(o[poly] + err1, q2), q);
}

} // namespace

void launch_tile_unsigned(const DeviceTables &t, const int64_t *samples,
                          int64_t *out, int count, int limbs, int moduli_id,
                          hipStream_t stream) {
  const int64_t total = (int64_t(count) * limbs) << t.log_n;
  hipLaunchKernelGGL(tile_unsigned_kernel, dim3(grid_for(total)),
                     dim3(kBlock), 0, stream, samples, out, total, limbs,
                     moduli_id, t);
}

void launch_tile_monomial(const DeviceTables &t, int64_t value, int64_t pos,
                          int64_t *out, int limbs, int moduli_id,
                          hipStream_t stream) {
  const int64_t total = int64_t(limbs) << t.log_n;
  hipLaunchKernelGGL(tile_monomial_kernel, dim3(grid_for(total)),
                     dim3(kBlock), 0, stream, out, total, moduli_id, value,
                     pos, t);
}

void launch_encrypt_finish(const DeviceTables &t, const int64_t *pt,
                           const int64_t *e0, const int64_t *e1, int64_t *out,
                           int count, int limbs, int moduli_id,
                           hipStream_t stream) {
  const int64_t total = (int64_t(count) * limbs) << t.log_n;
  hipLaunchKernelGGL(encrypt_finish_kernel, dim3(grid_for(total)),
                     dim3(kBlock), 0, stream, pt, e0, e1, out, total, limbs,
                     moduli_id, t);
}

} // namespace fhe
