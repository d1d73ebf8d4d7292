// Elementwise and key-switching kernels for the CKKS hot path (gfx950).
//
// Each kernel performs the reference's per-coefficient operation sequence in
// the same order (so lazy and signed intermediates reproduce bit for bit), but
// fuses the steps that the reference runs as separate full-polynomial passes
// into one pass over HBM.  Exact-arithmetic modular work: VALU only, no MFMA.
#include "kernels.h"

#include "moddown.h"
#include "modup.h"
#include "ntt_r16.h"
#include "params.h"

namespace fhe {

namespace {

constexpr int kMaxAlpha = CkksParams::kMaxAlpha;

__global__ __launch_bounds__(kBlock) void mont_mult_kernel(
    const int64_t *a, const int64_t *b, int64_t *out, int64_t total, int limbs,
    int moduli_id, DeviceTables t) {
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= total)
    return;
  const int p = moduli_id + int((idx >> t.log_n) % limbs);
  out[idx] = mont_mult(a[idx], b[idx], t.q[p], t.k[p]);
}

// out = mont(a, b) where b is one polynomial of `limbs` limbs applied to
// every polynomial of a (CKKSCore::multiply(ct, pt), src/ckks/core.cpp:
// 257-282, over a batch of ciphertexts and their polynomials)
__global__ __launch_bounds__(kBlock) void mont_mult_broadcast_kernel(
    const int64_t *a, const int64_t *b, int64_t *out, int64_t total,
    int limbs, int moduli_id, DeviceTables t) {
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= total)
    return;
  const int64_t period = int64_t(limbs) << t.log_n;
  const int64_t r = idx % period;
  const int p = moduli_id + int(r >> t.log_n);
  out[idx] = mont_mult(a[idx], b[r], t.q[p], t.k[p]);
}

// mont_mult_kernel and mont_mult_broadcast_kernel over kChunk coefficients
// of one limb per block (chunked_ok); BCAST: b is one polynomial of `limbs`
// limbs applied to every polynomial of a.
template <bool BCAST>
__global__ __launch_bounds__(kBlock) void mont_mult_chunk_kernel(
    const int64_t *a, const int64_t *b, int64_t *out, int limbs,
    int moduli_id, DeviceTables t) {
  const int64_t e0 = int64_t(blockIdx.x) * kChunk;
  const uint32_t flat = uint32_t(e0 >> t.log_n);
  const uint32_t l = flat % uint32_t(limbs);
  const int p = moduli_id + int(l);
  const uint64_t q = t.q[p], k = t.k[p];
  const int64_t idx = e0 + int64_t(threadIdx.x) * kPerThread;
  const int64_t bidx =
      BCAST ? (int64_t(l) << t.log_n) + (idx & ((int64_t(1) << t.log_n) - 1))
            : idx;
  int64_t x[kPerThread], y[kPerThread];
  load4(a + idx, x);
  load4(b + bidx, y);
#pragma unroll
  for (int g = 0; g < kPerThread; g++)
    x[g] = mont_mult(x[g], y[g], q, k);
  store4(out + idx, x);
}

__global__ __launch_bounds__(kBlock) void tensor_kernel(
    const int64_t *a, const int64_t *b, int64_t *trio, int64_t total,
    int limbs, int moduli_id, DeviceTables t) {
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= total)
    return;
  const int64_t poly = int64_t(limbs) << t.log_n;
  const int64_t c = idx / poly;
  const int64_t r = idx - c * poly;
  const int p = moduli_id + int(r >> t.log_n);
  const uint64_t q = t.q[p], k = t.k[p];
  const int64_t q2 = int64_t(t.q2[p]);
  const int64_t *ac = a + c * 2 * poly + r;
  const int64_t *bc = b + c * 2 * poly + r;
  int64_t *z = trio + c * 3 * poly + r;
  const int64_t a0 = ac[0], a1 = ac[poly], b0 = bc[0], b1 = bc[poly];
  z[0] = mont_mult_u(a0, b0, q, k);
  const int64_t z1 = mont_mult_u(a0, b1, q, k);
  const int64_t z3 = mont_mult_u(a1, b0, q, k);
  z[poly] = guard_u(guard_u(add64(z1, z3), q2), int64_t(q));
  z[2 * poly] = mont_mult_u(a1, b1, q, k);
}

// One thread per (ciphertext, coefficient): Garner digits of the partition
// in registers, then every limb of the extended basis.  With SIGNED false
// the partition's first prime is at most twice each of its other primes, so
// every difference and digit below is non-negative and under 2^62.
template <int MAXZ, bool SIGNED, bool LAZY>
__global__ __launch_bounds__(kBlock) void modup_kernel(
    const int64_t *x, int64_t x_stride, int64_t *ext, int64_t ext_stride,
    int count, int limbs, int moduli_id, int start, int size,
    const int64_t *cst, int ext_shoup_offset, DeviceTables t) {
  const int64_t n = int64_t(1) << t.log_n;
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= int64_t(count) * n)
    return;
  const int64_t c = idx >> t.log_n;
  const int64_t coef = idx & (n - 1);
  const int64_t *xs = x + c * x_stride + coef;
  const int ext_limbs = limbs + t.alpha;

  int64_t xv[MAXZ], d[MAXZ];
#pragma unroll
  for (int j = 0; j < MAXZ; j++)
    xv[j] = j < size ? xs[int64_t(start + j) * n] : 0;
  garner_digits<MAXZ, SIGNED>(xv, d, size, moduli_id + start, cst, t);
  const int64_t *ec = cst + CkksParams::kKsExtOffset;
  const uint64_t *es = reinterpret_cast<const uint64_t *>(cst + ext_shoup_offset);
  int64_t *out = ext + c * ext_stride + coef;
  for (int i = 0; i < ext_limbs; i++) {
    const int p = moduli_id + i;
    if (LAZY)
      out[int64_t(i) * n] = extend_limb_lazy<MAXZ>(
          d, size, i, ext_limbs, es, t.q[p],
          uint64_t(opaque_neg(int64_t(t.q[p]))),
          t.q[p] < kLazySmallBound);
    else
      out[int64_t(i) * n] = extend_limb<MAXZ, SIGNED>(
          d, size, i, ext_limbs, ec, t.q[p], t.k[p], int64_t(t.q2[p]));
  }
}

// One thread per (limb, coefficient); the key values are loaded once and
// applied to every ciphertext of the batch.
__global__ __launch_bounds__(kBlock) void key_inner_product_kernel(
    const int64_t *ext, int64_t ext_stride, const int64_t *key0,
    const int64_t *key1, int64_t *acc, int64_t acc_stride, int count,
    int ext_limbs, int moduli_id, int first, DeviceTables t) {
  const int64_t n = int64_t(1) << t.log_n;
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= int64_t(ext_limbs) * n)
    return;
  const int p = moduli_id + int(idx >> t.log_n);
  const uint64_t q = t.q[p], k = t.k[p];
  const int64_t q2 = int64_t(t.q2[p]);
  const int64_t k0 = key0[idx], k1 = key1[idx];
  const int64_t half = int64_t(ext_limbs) * n;
  for (int c = 0; c < count; c++) {
    const int64_t e = ext[c * ext_stride + idx];
    int64_t *a0 = acc + c * acc_stride + idx;
    int64_t *a1 = a0 + half;
    const int64_t p0 = mont_mult_u(e, k0, q, k);
    const int64_t p1 = mont_mult_u(e, k1, q, k);
    if (first) {
      *a0 = p0;
      *a1 = p1;
    } else {
      *a0 = guard_u(guard_u(add64(p0, *a0), q2), int64_t(q));
      *a1 = guard_u(guard_u(add64(p1, *a1), q2), int64_t(q));
    }
  }
}

// The addend of half h of ciphertext c is d + c * d_stride + h * limbs * N
// for h < d_polys and zero otherwise.  The limb loops run in groups of
// kMdGroup with the conventions of moddown_relin (moddown.h), whose
// restrict-qualified pointers the kernel's follow.

template <int MAXA, bool FULL, bool RESCALE>
__global__ __launch_bounds__(kBlock) void moddown_add_kernel(
    const int64_t *__restrict__ acc, const int64_t *__restrict__ dbase,
    int64_t d_stride, int d_polys, int64_t *__restrict__ out, int count,
    int limbs, int moduli_id, DeviceTables t) {
  const int64_t n = int64_t(1) << t.log_n;
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= int64_t(count) * 2 * n)
    return;
  const int64_t ch = idx >> t.log_n; // ciphertext * 2 + half
  const int64_t coef = idx & (n - 1);
  const int64_t c = ch >> 1, h = ch & 1;
  const int ext_limbs = limbs + t.alpha;
  const int64_t *a = acc + ch * ext_limbs * n + coef;
  // without an addend the loads read the accumulator's limbs and are
  // discarded, so that every load is unconditional
  const bool has_d = h < d_polys;
  const int64_t *d = has_d ? dbase + c * d_stride + h * limbs * n + coef : a;
  if (!RESCALE) {
    moddown_relin<MAXA, FULL>(a, d, has_d, out + ch * limbs * n + coef, limbs,
                              moduli_id, n, t);
    return;
  }

  ModDown<MAXA, FULL> md;
  md.alpha = t.alpha;
  md.limbs = limbs;
  md.moduli_id = moduli_id;
  md.prepare(a, n, t);

  // relinearize: y_i = reduce(d_i + ks_i) (src/ckks/core.cpp:736-737); a
  // canonical ks_i passes the two guards unchanged when there is no d
  auto relin = [&](int i, int64_t av, int64_t dv) {
    const int p = moduli_id + i;
    const int64_t ks = md.limb_lazy(av, i, t);
    const int64_t di = has_d ? dv : 0;
    return guard(guard(di + ks, int64_t(t.q2[p])), int64_t(t.q[p]));
  };

  // rescale (src/ckks/core_rescale.cpp:5-25): drop limb 0 with rounding;
  // limbs [1, limbs) in groups as in moddown_relin
  const int64_t y0 = relin(0, a[0], d[0]);
  const int64_t rounder = uint64_t(y0) > t.qh[moduli_id] ? 1 : 0;
  int64_t *o = out + ch * (limbs - 1) * n + coef;
  for (int i0 = 1; i0 < limbs; i0 += kMdGroup) {
    int64_t av[kMdGroup], dv[kMdGroup];
#pragma unroll
    for (int g = 0; g < kMdGroup; g++) {
      const int i = min(i0 + g, limbs - 1);
      av[g] = a[int64_t(i) * n];
      dv[g] = d[int64_t(i) * n];
    }
#pragma unroll
    for (int g = 0; g < kMdGroup; g++) {
      const int i = min(i0 + g, limbs - 1);
      const int64_t y = relin(i, av[g], dv[g]);
      const int p = moduli_id + i;
      const uint64_t q = t.q[p], k = t.k[p];
      const int64_t q2 = int64_t(t.q2[p]);
      int64_t v = y - y0;
      v = v >= 0 ? v : v + q2;
      v = guard(v, int64_t(q));
      // |v| < 2^62 (y and y0 canonical): the product with the constant in
      // its signed-operand form, the same value as mont_mult
      v = mont_mult_cs(t.rescale[int64_t(moduli_id) * t.prime_count + p], v,
                       q, k);
      v = guard(v + rounder, q2);
      o[int64_t(i - 1) * n] = guard(v, int64_t(q));
    }
  }
}

// ModDown of one extended polynomial per (poly, coefficient) into its
// chain limbs, canonical (the relinearizing kernel above without an addend).
template <int MAXA, bool FULL>
__global__ __launch_bounds__(kBlock) void moddown_polys_kernel(
    const int64_t *__restrict__ acc, int64_t *__restrict__ out,
    int64_t polys, int limbs, int moduli_id, DeviceTables t) {
  const int64_t n = int64_t(1) << t.log_n;
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= (polys << t.log_n))
    return;
  const int64_t poly = idx >> t.log_n;
  const int64_t coef = idx & (n - 1);
  const int64_t *a = acc + poly * (limbs + t.alpha) * n + coef;
  ModDown<MAXA, FULL> md;
  md.alpha = t.alpha;
  md.limbs = limbs;
  md.moduli_id = moduli_id;
  md.prepare(a, n, t);
  int64_t *o = out + poly * limbs * n + coef;
  for (int i0 = 0; i0 < limbs; i0 += kMdGroup) {
    int64_t av[kMdGroup];
#pragma unroll
    for (int g = 0; g < kMdGroup; g++)
      av[g] = a[int64_t(min(i0 + g, limbs - 1)) * n];
#pragma unroll
    for (int g = 0; g < kMdGroup; g++) {
      const int i = min(i0 + g, limbs - 1);
      o[int64_t(i) * n] = md.limb_lazy(av[g], i, t);
    }
  }
}

// The limbs go in groups of kMdGroup as in moddown_relin (moddown.h); the
// output never overlaps the input.
__global__ __launch_bounds__(kBlock) void rescale_kernel(
    const int64_t *__restrict__ in, int64_t *__restrict__ out, int polys,
    int limbs, int moduli_id, DeviceTables t) {
  const int64_t n = int64_t(1) << t.log_n;
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= int64_t(polys) * n)
    return;
  const int64_t poly = idx >> t.log_n;
  const int64_t coef = idx & (n - 1);
  const int64_t *x = in + poly * limbs * n + coef;
  int64_t *o = out + poly * (limbs - 1) * n + coef;
  const int64_t x0 = x[0];
  const int64_t rounder = uint64_t(x0) > t.qh[moduli_id] ? 1 : 0;
  for (int i0 = 1; i0 < limbs; i0 += kMdGroup) {
    int64_t xv[kMdGroup];
#pragma unroll
    for (int g = 0; g < kMdGroup; g++)
      xv[g] = x[int64_t(min(i0 + g, limbs - 1)) * n];
#pragma unroll
    for (int g = 0; g < kMdGroup; g++) {
      const int i = min(i0 + g, limbs - 1);
      const int p = moduli_id + i;
      const uint64_t q = t.q[p], k = t.k[p];
      const int64_t q2 = int64_t(t.q2[p]);
      int64_t v = xv[g] - x0;
      v = v >= 0 ? v : v + q2;
      v = guard(v, int64_t(q));
      v = mont_mult(v, t.rescale[int64_t(moduli_id) * t.prime_count + p], q,
                    k);
      v = guard(v + rounder, q2);
      o[int64_t(i - 1) * n] = guard(v, int64_t(q));
    }
  }
}

// Scalar multiply then rescale (CKKSCore::multiply(ct, double),
// src/ckks/core.cpp:196-226, and CKKSCore::level_down, :792-838):
// y_i = guard(mont(x_i, c_i), q_i) on every source limb, then
// rescale_once_ (src/ckks/core_rescale.cpp:5-25) dropping limb 0, and the
// first `skip` limbs of the rescaled result are dropped as well.
struct ScalarRescaleArgs {
  const int64_t *in;
  int64_t *out;
  int polys;
  int limbs; // source limbs
  int moduli_id;
  int skip;
  int64_t c[kMaxEwLimbs];
};

__global__ __launch_bounds__(kBlock) void scalar_rescale_kernel(
    ScalarRescaleArgs a, DeviceTables t) {
  const int64_t n = int64_t(1) << t.log_n;
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= int64_t(a.polys) * n)
    return;
  const int64_t poly = idx >> t.log_n;
  const int64_t coef = idx & (n - 1);
  const int64_t *x = a.in + poly * a.limbs * n + coef;
  const int out_limbs = a.limbs - 1 - a.skip;
  int64_t *o = a.out + poly * out_limbs * n + coef;
  auto entered = [&](int i, int64_t xi) {
    const int p = a.moduli_id + i;
    return guard(mont_mult(xi, a.c[i], t.q[p], t.k[p]), int64_t(t.q[p]));
  };
  const int64_t y0 = entered(0, x[0]);
  const int64_t rounder = uint64_t(y0) > t.qh[a.moduli_id] ? 1 : 0;
  // limbs in load groups as in moddown_relin (moddown.h)
  for (int i0 = 1 + a.skip; i0 < a.limbs; i0 += kMdGroup) {
    int64_t xv[kMdGroup];
#pragma unroll
    for (int g = 0; g < kMdGroup; g++)
      xv[g] = x[int64_t(min(i0 + g, a.limbs - 1)) * n];
#pragma unroll
    for (int g = 0; g < kMdGroup; g++) {
      const int i = min(i0 + g, a.limbs - 1);
      const int p = a.moduli_id + i;
      const uint64_t q = t.q[p], k = t.k[p];
      const int64_t q2 = int64_t(t.q2[p]);
      int64_t v = entered(i, xv[g]) - y0;
      v = v >= 0 ? v : v + q2;
      v = guard(v, int64_t(q));
      // |v| < 2^62 (entered values lie in (-q, q)): as in moddown_add_kernel
      v = mont_mult_cs(t.rescale[int64_t(a.moduli_id) * t.prime_count + p], v,
                       q, k);
      v = guard(v + rounder, q2);
      o[int64_t(i - 1 - a.skip) * n] = guard(v, int64_t(q));
    }
  }
}

// weighted_sum (src/ckks/engine_advanced.cpp:228-269) of texts at one level
// and form in one pass: term t is multiply(basis, weight) - the scalar
// product with its multiplier constants and the rescale of
// scalar_rescale_kernel - and the terms are summed in order with add's
// guard(guard(a + b, 2q), q), so every limb equals the sequence of calls.
struct WeightedSumArgs {
  const int64_t *in[kWsMaxTerms]; // each [polys][limbs][N]
  int64_t *out;                   // [polys][limbs - 1][N]
  int terms;
  int polys;
  int limbs;                      // source limbs
  int moduli_id;
  int64_t c[kWsMaxTerms][kWsMaxLimbs];
};

// TERMS = a.terms, so that the term loops are straight-line code whose
// loads issue together.
template <int TERMS>
__global__ __launch_bounds__(kBlock) void weighted_sum_kernel(
    WeightedSumArgs a, DeviceTables t) {
  const int64_t n = int64_t(1) << t.log_n;
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= int64_t(a.polys) * n)
    return;
  const int64_t poly = idx >> t.log_n;
  const int64_t coef = idx & (n - 1);
  const int64_t src = poly * a.limbs * n + coef;
  int64_t *o = a.out + poly * (a.limbs - 1) * n + coef;
  const uint64_t q0 = t.q[a.moduli_id], k0 = t.k[a.moduli_id];
  int64_t y0[TERMS];
  int rounder[TERMS];
#pragma unroll
  for (int x = 0; x < TERMS; x++) {
    y0[x] = guard(mont_mult(a.in[x][src], a.c[x][0], q0, k0), int64_t(q0));
    rounder[x] = uint64_t(y0[x]) > t.qh[a.moduli_id] ? 1 : 0;
  }
  for (int i = 1; i < a.limbs; i++) {
    const int p = a.moduli_id + i;
    const uint64_t q = t.q[p], k = t.k[p];
    const int64_t q2 = int64_t(t.q2[p]), qq = int64_t(q);
    const int64_t rs = t.rescale[int64_t(a.moduli_id) * t.prime_count + p];
    int64_t xv[TERMS];
#pragma unroll
    for (int x = 0; x < TERMS; x++)
      xv[x] = a.in[x][src + int64_t(i) * n];
    int64_t acc = 0;
#pragma unroll
    for (int x = 0; x < TERMS; x++) {
      int64_t v = guard(mont_mult(xv[x], a.c[x][i], q, k), qq) - y0[x];
      v = v >= 0 ? v : v + q2;
      v = guard(v, qq);
      v = mont_mult_cs(rs, v, q, k); // |v| < 2^62, as in scalar_rescale
      v = guard(guard(v + rounder[x], q2), qq);
      acc = x == 0 ? v : guard(guard(acc + v, q2), qq);
    }
    o[int64_t(i - 1) * n] = acc;
  }
}

} // namespace

void launch_weighted_sum(const DeviceTables &t,
                         const int64_t *const *inputs,
                         const int64_t *const *constants, int terms,
                         int64_t *out, int polys, int limbs, int moduli_id,
                         hipStream_t stream) {
  WeightedSumArgs a{};
  a.out = out;
  a.terms = terms;
  a.polys = polys;
  a.limbs = limbs;
  a.moduli_id = moduli_id;
  for (int x = 0; x < terms; x++) {
    a.in[x] = inputs[x];
    for (int i = 0; i < limbs; i++)
      a.c[x][i] = constants[x][i];
  }
  const dim3 grid(grid_for(int64_t(polys) << t.log_n));
  switch (terms) {
#define FHE_WS(T)                                                              \
  case T:                                                                      \
    hipLaunchKernelGGL(weighted_sum_kernel<T>, grid, dim3(kBlock), 0, stream, \
                       a, t);                                                  \
    break;
    FHE_WS(1)
    FHE_WS(2)
    FHE_WS(3)
    FHE_WS(4)
    FHE_WS(5)
    FHE_WS(6)
    FHE_WS(7)
    FHE_WS(8)
#undef FHE_WS
  default:
    throw std::invalid_argument("weighted sum: 1 to 8 terms per launch");
  }
}

void launch_scalar_rescale(const DeviceTables &t, const int64_t *in,
                           int64_t *out, int polys, int limbs, int moduli_id,
                           const int64_t *constants, int skip,
                           hipStream_t stream) {
  ScalarRescaleArgs a{};
  a.in = in;
  a.out = out;
  a.polys = polys;
  a.limbs = limbs;
  a.moduli_id = moduli_id;
  a.skip = skip;
  for (int l = 0; l < limbs; l++)
    a.c[l] = constants[l];
  const dim3 grid(grid_for(int64_t(polys) << t.log_n));
  hipLaunchKernelGGL(scalar_rescale_kernel, grid, dim3(kBlock), 0, stream, a,
                     t);
}

void launch_mont_mult(const DeviceTables &t, const int64_t *a,
                      const int64_t *b, int64_t *out, int64_t total_limbs,
                      int limbs, int moduli_id, hipStream_t stream) {
  const int64_t total = total_limbs << t.log_n;
  if (chunked_ok(t.log_n, {a, b, out})) {
    hipLaunchKernelGGL(mont_mult_chunk_kernel<false>,
                       dim3(unsigned(total / kChunk)), dim3(kBlock), 0, stream,
                       a, b, out, limbs, moduli_id, t);
    return;
  }
  hipLaunchKernelGGL(mont_mult_kernel, dim3(grid_for(total)), dim3(kBlock), 0,
                     stream, a, b, out, total, limbs, moduli_id, t);
}

void launch_mont_mult_broadcast(const DeviceTables &t, const int64_t *a,
                                const int64_t *b, int64_t *out,
                                int64_t total_limbs, int limbs, int moduli_id,
                                hipStream_t stream) {
  const int64_t total = total_limbs << t.log_n;
  if (chunked_ok(t.log_n, {a, b, out})) {
    hipLaunchKernelGGL(mont_mult_chunk_kernel<true>,
                       dim3(unsigned(total / kChunk)), dim3(kBlock), 0, stream,
                       a, b, out, limbs, moduli_id, t);
    return;
  }
  hipLaunchKernelGGL(mont_mult_broadcast_kernel, dim3(grid_for(total)),
                     dim3(kBlock), 0, stream, a, b, out, total, limbs,
                     moduli_id, t);
}

void launch_tensor(const DeviceTables &t, const int64_t *a, const int64_t *b,
                   int64_t *trio, int count, int limbs, int moduli_id,
                   hipStream_t stream) {
  const int64_t total = (int64_t(count) * limbs) << t.log_n;
  hipLaunchKernelGGL(tensor_kernel, dim3(grid_for(total)), dim3(kBlock), 0,
                     stream, a, b, trio, total, limbs, moduli_id, t);
}

void launch_modup(const DeviceTables &t, const int64_t *x, int64_t x_stride,
                  int64_t *ext, int64_t ext_stride, int count, int limbs,
                  int moduli_id, int start, int size, const int64_t *constants,
                  int ext_shoup_offset, bool non_negative, bool lazy,
                  hipStream_t stream) {
  const dim3 grid(grid_for(int64_t(count) << t.log_n));
#define FHE_MODUP(W, S, L)                                                     \
  hipLaunchKernelGGL((modup_kernel<W, S, L>), grid, dim3(kBlock), 0, stream,  \
                     x, x_stride, ext, ext_stride, count, limbs, moduli_id,    \
                     start, size, constants, ext_shoup_offset, t)
#define FHE_MODUP_W(W)                                                         \
  do {                                                                         \
    if (non_negative && lazy)                                                  \
      FHE_MODUP(W, false, true);                                               \
    else if (non_negative)                                                     \
      FHE_MODUP(W, false, false);                                              \
    else                                                                       \
      FHE_MODUP(W, true, false);                                               \
  } while (0)
  if (size <= 1)
    FHE_MODUP_W(1);
  else if (size <= 2)
    FHE_MODUP_W(2);
  else if (size <= 4)
    FHE_MODUP_W(4);
  else if (size <= 8)
    FHE_MODUP_W(8);
  else
    FHE_MODUP_W(16);
#undef FHE_MODUP_W
#undef FHE_MODUP
}

void launch_key_inner_product(const DeviceTables &t, const int64_t *ext,
                              int64_t ext_stride, const int64_t *key0,
                              const int64_t *key1, int64_t *acc,
                              int64_t acc_stride, int count, int ext_limbs,
                              int moduli_id, bool first, hipStream_t stream) {
  const dim3 grid(grid_for(int64_t(ext_limbs) << t.log_n));
  hipLaunchKernelGGL(key_inner_product_kernel, grid, dim3(kBlock), 0, stream,
                     ext, ext_stride, key0, key1, acc, acc_stride, count,
                     ext_limbs, moduli_id, first ? 1 : 0, t);
}

template <bool RESCALE>
static void launch_moddown(const DeviceTables &t, const int64_t *acc,
                           const int64_t *d, int64_t d_stride, int d_polys,
                           int64_t *out, int count, int limbs, int moduli_id,
                           hipStream_t stream) {
  const dim3 grid(grid_for((int64_t(count) * 2) << t.log_n));
#define FHE_MODDOWN(W)                                                         \
  do {                                                                         \
    if (t.alpha == W)                                                          \
      hipLaunchKernelGGL((moddown_add_kernel<W, true, RESCALE>), grid,        \
                         dim3(kBlock), 0, stream, acc, d, d_stride, d_polys,   \
                         out, count, limbs, moduli_id, t);                     \
    else                                                                       \
      hipLaunchKernelGGL((moddown_add_kernel<W, false, RESCALE>), grid,       \
                         dim3(kBlock), 0, stream, acc, d, d_stride, d_polys,   \
                         out, count, limbs, moduli_id, t);                     \
  } while (0)
  if (t.alpha <= 1)
    FHE_MODDOWN(1);
  else if (t.alpha <= 2)
    FHE_MODDOWN(2);
  else if (t.alpha <= 4)
    FHE_MODDOWN(4);
  else if (t.alpha <= 8)
    FHE_MODDOWN(8);
  else
    FHE_MODDOWN(16);
#undef FHE_MODDOWN
}

void launch_moddown_relin_rescale(const DeviceTables &t, const int64_t *acc,
                                  const int64_t *trio, int64_t *out, int count,
                                  int limbs, int moduli_id,
                                  hipStream_t stream) {
  launch_moddown<true>(t, acc, trio, int64_t(3) * limbs << t.log_n, 2, out,
                       count, limbs, moduli_id, stream);
}

void launch_moddown_relin(const DeviceTables &t, const int64_t *acc,
                          const int64_t *trio, int64_t *out, int count,
                          int limbs, int moduli_id, hipStream_t stream) {
  launch_moddown<false>(t, acc, trio, int64_t(3) * limbs << t.log_n, 2, out,
                        count, limbs, moduli_id, stream);
}

void launch_moddown_polys(const DeviceTables &t, const int64_t *acc,
                          int64_t *out, int64_t polys, int limbs,
                          int moduli_id, hipStream_t stream) {
  const dim3 grid(grid_for(polys << t.log_n));
#define FHE_MODDOWN_POLYS(W)                                                   \
  do {                                                                         \
    if (t.alpha == W)                                                          \
      hipLaunchKernelGGL((moddown_polys_kernel<W, true>), grid, dim3(kBlock), \
                         0, stream, acc, out, polys, limbs, moduli_id, t);     \
    else                                                                       \
      hipLaunchKernelGGL((moddown_polys_kernel<W, false>), grid,              \
                         dim3(kBlock), 0, stream, acc, out, polys, limbs,      \
                         moduli_id, t);                                        \
  } while (0)
  if (t.alpha <= 1)
    FHE_MODDOWN_POLYS(1);
  else if (t.alpha <= 2)
    FHE_MODDOWN_POLYS(2);
  else if (t.alpha <= 4)
    FHE_MODDOWN_POLYS(4);
  else if (t.alpha <= 8)
    FHE_MODDOWN_POLYS(8);
  else
    FHE_MODDOWN_POLYS(16);
#undef FHE_MODDOWN_POLYS
}

void launch_moddown_add(const DeviceTables &t, const int64_t *acc,
                        const int64_t *d, int64_t d_stride, int d_polys,
                        int64_t *out, int count, int limbs, int moduli_id,
                        hipStream_t stream) {
  launch_moddown<false>(t, acc, d, d_stride, d_polys, out, count, limbs,
                        moduli_id, stream);
}

void launch_rescale(const DeviceTables &t, const int64_t *in, int64_t *out,
                    int polys, int limbs, int moduli_id, hipStream_t stream) {
  const dim3 grid(grid_for(int64_t(polys) << t.log_n));
  hipLaunchKernelGGL(rescale_kernel, grid, dim3(kBlock), 0, stream, in, out,
                     polys, limbs, moduli_id, t);
}

} // namespace fhe
