// Device-side tables shared by every kernel, and launch helpers.
#pragma once

#include <cstdint>
#include <initializer_list>

#include <hip/hip_runtime.h>

#include "modarith.h"

namespace fhe {

// Per-prime tables indexed by ABSOLUTE prime index in the chain (limb l of an
// object at moduli_id m uses prime m + l).  All pointers are device memory
// owned by the context.
struct DeviceTables {
  const uint64_t *q;
  const uint64_t *k;
  const uint64_t *q2;
  const uint64_t *qh;
  const int64_t *r2;
  const int64_t *ninv;
  const int64_t *root;          // [prime][N]
  const int64_t *root_inverse;  // [prime][N]
  // standard-form twiddles with Shoup quotients, [prime][N][2]
  const uint64_t *root_shoup;
  const uint64_t *root_inverse_shoup;
  const int64_t *p_inverse;     // [alpha][prime_count]
  const uint64_t *p_inverse_shoup; // [alpha][prime_count][2] standard form
  const uint64_t *q_top;        // largest multiple of q below 2^63
  const int64_t *rescale;       // [prime_count][prime_count]
  // HOST copy of the primes (for launch decisions on the host; kernels
  // never read it)
  const uint64_t *host_q;
  // launch choice (host): one launch per arithmetic variant instead of a
  // per-block choice inside one kernel
  int split_variants;
  int prime_count;
  int log_n;
  int alpha;
};

// Post-operation applied by the final inverse-NTT pass.
enum InttPost : int {
  kInttRaw = 0,       // plain INTT output, lazy in [0, 2q)
  kInttExit = 1,      // * N^-1, leave Montgomery form (to_intt)
  kInttExitReduce = 2 // as kInttExit, then reduce [0, 2q) -> [0, q)
};

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t work, int per_block = kBlock) {
  return unsigned((work + per_block - 1) / per_block);
}

// Chunked per-coefficient kernels: a block covers kChunk consecutive
// coefficients of one limb, kPerThread per thread (two 16-byte accesses per
// operand), so that the limb, its prime and its constants are uniform in
// the block.  They take limbs of N >= kChunk coefficients and operands at
// 16-byte aligned addresses (chunked_ok); the one-coefficient-per-thread
// kernels serve the rest.
constexpr int kPerThread = 4;
constexpr int64_t kChunk = int64_t(kBlock) * kPerThread;

inline bool chunked_ok(int log_n, std::initializer_list<const void *> ptrs) {
  if ((int64_t(1) << log_n) % kChunk != 0)
    return false;
  for (const void *p : ptrs)
    if (reinterpret_cast<uintptr_t>(p) % 16 != 0)
      return false;
  return true;
}

#if defined(__HIPCC__)
// The thread's index in its block as a value that the compiler cannot take
// for invariant across the iterations of a persistent loop: otherwise it
// keeps every per-thread offset derived from the index (the LDS and global
// offsets of a tile's sixteen points) live across the loop, at the cost of
// occupancy.  An empty asm statement that claims to modify the value.
__device__ __forceinline__ int opaque_tid() {
  int tid = int(threadIdx.x);
  asm volatile("" : "+v"(tid));
  return tid;
}

// -q for guard_neg, behind an empty asm statement that claims to modify it,
// so that the compiler cannot fold x + (-q) back into a subtraction.
__device__ __forceinline__ int64_t opaque_neg(int64_t q) {
  int64_t n = -q;
  asm("" : "+v"(n));
  return n;
}

typedef int64_t i64x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void load4(const int64_t *p, int64_t (&v)[4]) {
  const i64x2 a = reinterpret_cast<const i64x2 *>(p)[0];
  const i64x2 b = reinterpret_cast<const i64x2 *>(p)[1];
  v[0] = a.x;
  v[1] = a.y;
  v[2] = b.x;
  v[3] = b.y;
}

__device__ __forceinline__ void store4(int64_t *p, const int64_t (&v)[4]) {
  i64x2 a, b;
  a.x = v[0];
  a.y = v[1];
  b.x = v[2];
  b.y = v[3];
  reinterpret_cast<i64x2 *>(p)[0] = a;
  reinterpret_cast<i64x2 *>(p)[1] = b;
}
#endif

} // namespace fhe
