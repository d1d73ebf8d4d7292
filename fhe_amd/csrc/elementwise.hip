// Per-coefficient Montgomery primitives (the reference's unit and compound
// limb operations, src/montgomery/unit.cpp:5-86, src/montgomery/compound.cpp:
// 6-109, and add_scalar, src/operation/scalar.cpp:5-11) over a batch of
// [polys][limbs][N] limbs in one launch.
//
// The operation code is uniform across the grid, so the switch costs one
// scalar branch per element.  Operands may be negative where the reference's
// can (mont_sub_enter feeds signed values to its product), so the signed
// Montgomery product is used throughout; these passes are HBM-bound and the
// extra multiply-adds are free.  Every formula evaluates the reference's
// expression in the same order, which keeps lazy representatives identical.
#include <algorithm>

#include "kernels.h"

namespace fhe {

namespace {

struct EwArgs {
  const int64_t *x1;
  const int64_t *x2;
  int64_t *y;
  int64_t total;
  int limbs;
  int moduli_id;
  int op;
  int copy_period;        // > 0: polynomials p % copy_period != 0 are copied
  int64_t c[kMaxEwLimbs]; // per-limb constant, limb l of every polynomial
};

__device__ __forceinline__ int64_t ew_apply(int op, int64_t a, int64_t b,
                                            int64_t y_old, int64_t c,
                                            uint64_t q, uint64_t k, int64_t q2,
                                            int64_t qh, bool first_coeff) {
  const int64_t qq = int64_t(q);
  switch (op) {
  case kEwMontMult:
    return mont_mult(a, b, q, k);
  case kEwMontEnter:
    return mont_mult(a, c, q, k);
  case kEwReduce:
    return guard(a, qq);
  // make_signed and get_rounder compare with an unsigned q/2: a negative
  // operand counts as large there
  case kEwMakeSigned:
    return uint64_t(a) <= uint64_t(qh) ? a : a - qq;
  case kEwNegate:
    return -a;
  case kEwAdd:
    return guard(a + b, q2);
  case kEwSubtract: {
    const int64_t d = a - b;
    return d >= 0 ? d : d + q2;
  }
  case kEwEnterAdd:
    return guard(mont_mult(a, c, q, k) + b, q2);
  case kEwEnterAddReduce:
    return guard(guard(mont_mult(a, c, q, k) + b, q2), qq);
  case kEwMultAdd:
    return guard(mont_mult(a, b, q, k) + y_old, q2);
  case kEwSubEnterReduce:
    return guard(mont_mult(guard(a + q2 - b, q2), c, q, k), qq);
  case kEwSubEnter:
    return mont_mult(guard(a + q2 - b, q2), c, q, k);
  case kEwEnterExit:
    return mont_exit(mont_mult(a, c, q, k), q, k);
  case kEwEnterExitReduce:
    return guard(mont_exit(mont_mult(a, c, q, k), q, k), qq);
  case kEwAddReduce:
    return guard(guard(a + b, q2), qq);
  case kEwSubReduce:
    return guard(guard(a + q2 - b, q2), qq);
  case kEwMakeUnsignedReduce:
    return a < 0 ? a + qq : guard(a, qq);
  case kEwNegateMakeUnsigned:
    return qq - a;
  case kEwAddScalar:
    return first_coeff ? a + c : a;
  case kEwRounder:
    return uint64_t(a) > uint64_t(qh) ? 1 : 0;
  case kEwEnterReduce:
    return guard(mont_mult(a, c, q, k), qq);
  case kEwAddScalarReduce:
    return guard(first_coeff ? a + c : a, qq);
  default:
    return 0;
  }
}

__global__ __launch_bounds__(kBlock) void elementwise_kernel(EwArgs a,
                                                             DeviceTables t) {
  const int64_t stride = int64_t(gridDim.x) * kBlock;
  for (int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
       idx < a.total; idx += stride) {
    const int64_t flat = idx >> t.log_n;
    const int l = int(flat % a.limbs);
    if (a.copy_period > 0 && (flat / a.limbs) % a.copy_period != 0) {
      a.y[idx] = a.x1[idx];
      continue;
    }
    const int p = a.moduli_id + l;
    const bool first = (idx & ((int64_t(1) << t.log_n) - 1)) == 0;
    const int64_t x2 = a.x2 ? a.x2[idx] : 0;
    const int64_t y_old = a.op == kEwMultAdd ? a.y[idx] : 0;
    a.y[idx] = ew_apply(a.op, a.x1[idx], x2, y_old, a.c[l], t.q[p], t.k[p],
                        int64_t(t.q2[p]), int64_t(t.qh[p]), first);
  }
}

// The same over kChunk coefficients of one limb per block (chunked_ok).
__global__ __launch_bounds__(kBlock) void elementwise_chunk_kernel(
    EwArgs a, DeviceTables t) {
  const int64_t e0 = int64_t(blockIdx.x) * kChunk;
  // block-uniform: the limb, its prime and constants
  const uint32_t flat = uint32_t(e0 >> t.log_n);
  const int l = int(flat % uint32_t(a.limbs));
  const int64_t idx = e0 + int64_t(threadIdx.x) * kPerThread;
  int64_t x1[kPerThread];
  load4(a.x1 + idx, x1);
  if (a.copy_period > 0 &&
      (flat / uint32_t(a.limbs)) % uint32_t(a.copy_period) != 0) {
    store4(a.y + idx, x1);
    return;
  }
  const int p = a.moduli_id + l;
  const uint64_t q = t.q[p], k = t.k[p];
  const int64_t q2 = int64_t(t.q2[p]), qh = int64_t(t.qh[p]);
  const int64_t c = a.c[l];
  int64_t x2[kPerThread] = {}, y_old[kPerThread] = {};
  int64_t y[kPerThread];
  if (a.x2)
    load4(a.x2 + idx, x2);
  if (a.op == kEwMultAdd)
    load4(a.y + idx, y_old);
  const bool first = (idx & ((int64_t(1) << t.log_n) - 1)) == 0;
#pragma unroll
  for (int g = 0; g < kPerThread; g++)
    y[g] = ew_apply(a.op, x1[g], x2[g], y_old[g], c, q, k, q2, qh,
                    first && g == 0);
  store4(a.y + idx, y);
}

} // namespace

bool ew_needs_x2(int op) {
  switch (op) {
  case kEwMontMult:
  case kEwAdd:
  case kEwSubtract:
  case kEwEnterAdd:
  case kEwEnterAddReduce:
  case kEwMultAdd:
  case kEwSubEnterReduce:
  case kEwSubEnter:
  case kEwAddReduce:
  case kEwSubReduce:
    return true;
  default:
    return false;
  }
}

bool ew_needs_constants(int op) {
  switch (op) {
  case kEwMontEnter:
  case kEwEnterAdd:
  case kEwEnterAddReduce:
  case kEwSubEnterReduce:
  case kEwSubEnter:
  case kEwEnterExit:
  case kEwEnterExitReduce:
  case kEwAddScalar:
  case kEwEnterReduce:
  case kEwAddScalarReduce:
    return true;
  default:
    return false;
  }
}

void launch_elementwise(const DeviceTables &t, int op, const int64_t *x1,
                        const int64_t *x2, int64_t *y, int64_t total_limbs,
                        int limbs, int moduli_id, const int64_t *constants,
                        hipStream_t stream, int copy_period) {
  EwArgs a{};
  a.copy_period = copy_period;
  a.x1 = x1;
  a.x2 = x2;
  a.y = y;
  a.total = total_limbs << t.log_n;
  a.limbs = limbs;
  a.moduli_id = moduli_id;
  a.op = op;
  for (int l = 0; l < limbs && constants; l++)
    a.c[l] = constants[l];
  if (chunked_ok(t.log_n, {x1, x2, y})) {
    hipLaunchKernelGGL(elementwise_chunk_kernel,
                       dim3(unsigned(a.total / kChunk)), dim3(kBlock), 0,
                       stream, a, t);
    return;
  }
  // enough blocks to fill 256 CUs several times over; each thread strides
  const int64_t blocks = std::min<int64_t>(grid_for(a.total), 256 * 64);
  hipLaunchKernelGGL(elementwise_kernel, dim3(unsigned(blocks)), dim3(kBlock),
                     0, stream, a, t);
}

} // namespace fhe
