// Persistent work-list kernels for batches at uneven levels (uneven.h).
//
// Each kernel is one stage of the key-switching pipeline of the batched
// kernels (ntt.hip, ks_fused.hip, ckks_kernels.hip, automorphism.hip) with
// the same per-coefficient arithmetic - the item functions of ntt_r16.h,
// ks_tile.h, modup.h and moddown.h - but its blocks take their work units
// from a list instead of a (ciphertext, limb, tile) grid of one moduli_id.
#include "ks_tile.h"
#include "moddown.h"
#include "uneven.h"

namespace fhe {

namespace {

// The units [w, end) of this block, stride `step`.  With list.xcd and a
// grid that is a multiple of 8, XCD x (blocks x, x + 8, ...) walks the x-th
// eighth of the list.
struct UnitWalk {
  int w, end, step;
  __device__ explicit UnitWalk(const UnitList &l) {
    const int g = int(gridDim.x);
    const int b = int(blockIdx.x);
    if (l.xcd && g >= 8 && (g & 7) == 0) {
      const int chunk = (l.count + 7) >> 3;
      const int x = b & 7;
      w = x * chunk + (b >> 3);
      end = min(l.count, (x + 1) * chunk);
      step = g >> 3;
    } else {
      w = b;
      end = l.count;
      step = g;
    }
  }
};

unsigned grid_of(int count, int cap) {
  return unsigned(cap > 0 && cap < count ? cap : count);
}

// ---------------------------------------------------------------------------
struct UnevenNttArgs {
  UnitList list;
  const UnevenCt *cts;
  UnevenPtrs src, dst;
  const int64_t *pre;
  int ext_stride, b, post, last;
};

template <bool INV, bool COLS, int A>
__global__ __launch_bounds__(kBlock) void uneven_ntt_kernel(UnevenNttArgs a,
                                                            DeviceTables t) {
  __shared__ int64_t lds[kR16Lds];
  for (UnitWalk u(a.list); u.w < u.end; u.w += u.step) {
    const uint32_t unit = a.list.units[u.w];
    const int c = unit_ct(unit), l = unit_limb(unit);
    const UnevenCt &ct = a.cts[c];
    const int stride = a.ext_stride ? ct.ext : ct.limbs;
    const int64_t off = (int64_t(unit_a(unit)) * stride + l) << t.log_n;
    r16_limb_pass<INV, COLS, A, true>(a.src.p[c] + off,
                                const_cast<int64_t *>(a.dst.p[c]) + off,
                                ct.moduli_id + l, unit_tile(unit), a.b, a.pre,
                                a.post, a.last != 0, t, lds);
  }
}

// ---------------------------------------------------------------------------
struct UnevenDigitsArgs {
  UnitList list;
  const UnevenCt *cts;
  UnevenPtrs x, digits;
  int maxz;
};

template <int MAXZ>
__global__ __launch_bounds__(kBlock) void uneven_digits_kernel(
    UnevenDigitsArgs a, DeviceTables t) {
  for (UnitWalk u(a.list); u.w < u.end; u.w += u.step) {
    const uint32_t unit = a.list.units[u.w];
    const int c = unit_ct(unit), part = unit_a(unit);
    const UnevenCt &ct = a.cts[c];
    const int64_t coef = int64_t(unit_tile(unit)) * kBlock + threadIdx.x;
    const int start = ct.start[part], size = ct.size[part];
    const int64_t *xs = a.x.p[c] + coef;
    const int64_t *cst = ct.cst + int64_t(part) * ct.cst_stride;
    // every load unconditional (limbs past the partition's size reload its
    // last limb and are unused), so that they issue together
    int64_t xv[MAXZ], d[MAXZ];
#pragma unroll
    for (int j = 0; j < MAXZ; j++) {
      const int64_t v = xs[int64_t(start + min(j, size - 1)) << t.log_n];
      xv[j] = j < size ? v : 0;
    }
    if ((ct.signed_mask >> part) & 1)
      garner_digits<MAXZ, true>(xv, d, size, ct.moduli_id + start, cst, t);
    else
      garner_digits<MAXZ, false>(xv, d, size, ct.moduli_id + start, cst, t);
    int64_t *out = const_cast<int64_t *>(a.digits.p[c]) +
                   (int64_t(part) * a.maxz << t.log_n) + coef;
#pragma unroll
    for (int j = 0; j < MAXZ; j++)
      if (j < size)
        out[int64_t(j) << t.log_n] = d[j];
  }
}

// ---------------------------------------------------------------------------
struct UnevenExtArgs {
  UnitList list;
  const UnevenCt *cts;
  UnevenPtrs digits, ext;
  int maxz;
};

template <int A, int MAXZ, bool SIGNED>
__global__ __launch_bounds__(kBlock) void uneven_extend_kernel(
    UnevenExtArgs a, DeviceTables t) {
  __shared__ uint32_t lds[kR16LdsWords];
  for (UnitWalk u(a.list); u.w < u.end; u.w += u.step) {
    const uint32_t unit = a.list.units[u.w];
    const int c = unit_ct(unit), part = unit_a(unit);
    const UnevenCt &ct = a.cts[c];
    ExtItem it;
    it.digits = a.digits.p[c] + (int64_t(part) * a.maxz << t.log_n);
    it.cst = ct.cst + int64_t(part) * ct.cst_stride;
    it.ext = const_cast<int64_t *>(a.ext.p[c]) +
             (int64_t(part) * ct.ext << t.log_n);
    it.size = ct.size[part];
    it.ext_limbs = ct.ext;
    it.moduli_id = ct.moduli_id;
    it.ext_shoup_offset = ct.ext_shoup_offset;
    ext_item_tile<A, MAXZ, SIGNED, true>(it, t, lds, unit_tile(unit),
                                   unit_limb(unit));
  }
}

// ---------------------------------------------------------------------------
struct UnevenKsArgs {
  UnitList list;
  const UnevenCt *cts;
  UnevenPtrs ext, own, acc;
  const int64_t *key;
  int key_chain, key_moduli_id, lazy_acc;
};

// Addressing of one ciphertext of an uneven batch for ks_item_tile.
struct UnevenKsAddr {
  const UnevenCt &ct;
  const int64_t *ext_base, *own_base, *key;
  int64_t *acc_base;
  int log_n, key_offset;
  int parts, ext_limbs, moduli_id, key_chain, lazy_acc;
  __device__ UnevenKsAddr(const UnevenKsArgs &a, int c, int ln)
      : ct(a.cts[c]), ext_base(a.ext.p[c]), own_base(a.own.p[c]), key(a.key),
        acc_base(const_cast<int64_t *>(a.acc.p[c])), log_n(ln),
        key_offset(a.cts[c].moduli_id - a.key_moduli_id),
        parts(a.cts[c].parts), ext_limbs(a.cts[c].ext),
        moduli_id(a.cts[c].moduli_id), key_chain(a.key_chain),
        lazy_acc(a.lazy_acc) {}
  __device__ bool is_own(int pp, int i) const {
    return own_base != nullptr && i >= ct.start[pp] &&
           i < ct.start[pp] + ct.size[pp];
  }
  __device__ const int64_t *own(int i, int64_t off) const {
    return own_base + (int64_t(i) << log_n) + off;
  }
  __device__ const int64_t *ext(int pp, int i, int64_t off) const {
    return ext_base + ((int64_t(pp) * ext_limbs + i) << log_n) + off;
  }
  __device__ const int64_t *key_tile(int pp, int i, int64_t off) const {
    const int kp = ct.key_partition0 + pp;
    return key + ((int64_t(2 * kp) * key_chain + key_offset + i) << log_n) +
           off;
  }
  __device__ int64_t *acc(int i, int64_t off) const {
    return acc_base + (int64_t(i) << log_n) + off;
  }
};

template <int A, bool WIDE>
__global__ __launch_bounds__(kBlock) void uneven_ks_kernel(UnevenKsArgs a,
                                                           DeviceTables t) {
  __shared__ int64_t lds[kR16Lds];
  for (UnitWalk u(a.list); u.w < u.end; u.w += u.step) {
    const uint32_t unit = a.list.units[u.w];
    const UnevenKsAddr ad(a, unit_ct(unit), t.log_n);
    ks_item_tile<A, false, WIDE, UnevenKsAddr, true>(
        ad, t, lds, unit_tile(unit), unit_limb(unit));
  }
}

// ---------------------------------------------------------------------------
struct UnevenModDownArgs {
  UnitList list;
  const UnevenCt *cts;
  UnevenPtrs acc, d, out;
  int d_polys;
};

template <int MAXA, bool FULL>
__global__ __launch_bounds__(kBlock) void uneven_moddown_kernel(
    UnevenModDownArgs a, DeviceTables t) {
  const int64_t n = int64_t(1) << t.log_n;
  for (UnitWalk u(a.list); u.w < u.end; u.w += u.step) {
    const uint32_t unit = a.list.units[u.w];
    const int c = unit_ct(unit), h = unit_a(unit);
    const UnevenCt &ct = a.cts[c];
    const int limbs = ct.limbs;
    const int64_t coef = int64_t(unit_tile(unit)) * kBlock + threadIdx.x;
    const int64_t *acc = a.acc.p[c] + (int64_t(h) * ct.ext << t.log_n) + coef;
    // as moddown_add_kernel (no addend: the accumulator's limbs are read
    // in its place and discarded)
    const bool has_d = h < a.d_polys;
    const int64_t *d =
        has_d ? a.d.p[c] + (int64_t(h) * limbs << t.log_n) + coef : acc;
    int64_t *o = const_cast<int64_t *>(a.out.p[c]) +
                 (int64_t(h) * limbs << t.log_n) + coef;
    moddown_relin<MAXA, FULL>(acc, d, has_d, o, limbs, ct.moduli_id, n, t);
  }
}

// ---------------------------------------------------------------------------
struct UnevenAutArgs {
  UnitList list;
  const UnevenCt *cts;
  UnevenPtrs in, out;
  uint32_t h; // g^-1 mod 2N
};

constexpr int kAutChunk = 4 * kBlock;

// out[o] = +-in[h o mod 2N] with make_unsigned_reduce_2q_to_q, as
// automorphism_kernel (automorphism.hip).
__global__ __launch_bounds__(kBlock) void uneven_automorphism_kernel(
    UnevenAutArgs a, DeviceTables t) {
  const uint32_t n = 1u << t.log_n;
  for (UnitWalk u(a.list); u.w < u.end; u.w += u.step) {
    const uint32_t unit = a.list.units[u.w];
    const int c = unit_ct(unit), l = unit_limb(unit);
    const UnevenCt &ct = a.cts[c];
    const int64_t off = (int64_t(unit_a(unit)) * ct.limbs + l) << t.log_n;
    const int64_t *in = a.in.p[c] + off;
    int64_t *out = const_cast<int64_t *>(a.out.p[c]) + off;
    const int64_t q = int64_t(t.q[ct.moduli_id + l]);
    // the four gathers issue before the stores (out never overlaps in)
    uint32_t o[4], s[4];
    int64_t x[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
      o[e] = uint32_t(unit_tile(unit)) * kAutChunk + e * kBlock + threadIdx.x;
      s[e] = (a.h * o[e]) & (2 * n - 1);
      x[e] = in[s[e] & (n - 1)];
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int64_t y = (s[e] & n) ? -x[e] : x[e];
      out[o[e]] = y < 0 ? y + q : guard(y, q);
    }
  }
}

// g^-1 mod 2^32 for odd g (Newton: each step doubles the correct bits).
uint32_t odd_inverse32(uint32_t g) {
  uint32_t x = g;
  for (int i = 0; i < 4; i++)
    x *= 2u - g * x;
  return x;
}

} // namespace

void launch_uneven_ntt(const DeviceTables &t, const UnevenCt *cts,
                       const UnitList &list, const UnevenPtrs &src,
                       const UnevenPtrs &dst, int ext_stride, bool inverse,
                       int b, const int64_t *pre, int post, bool last,
                       int variant, int cap, hipStream_t stream) {
  if (list.count == 0)
    return;
  UnevenNttArgs a{list, cts, src, dst, pre, ext_stride, b, post, last ? 1 : 0};
  const dim3 grid(grid_of(list.count, cap));
  const bool cols = b > 0;
#define FHE_UNTT(I, C, A)                                                      \
  hipLaunchKernelGGL((uneven_ntt_kernel<I, C, A>), grid, dim3(kBlock), 0,      \
                     stream, a, t)
#define FHE_UNTT_A(I, C)                                                       \
  do {                                                                         \
    if (variant == kExact)                                                     \
      FHE_UNTT(I, C, kExact);                                                  \
    else if (variant == kLazySmall)                                            \
      FHE_UNTT(I, C, kLazySmall);                                              \
    else                                                                       \
      FHE_UNTT(I, C, kLazyLarge);                                              \
  } while (0)
  if (inverse) {
    if (cols)
      FHE_UNTT_A(true, true);
    else
      FHE_UNTT_A(true, false);
  } else {
    if (cols)
      FHE_UNTT_A(false, true);
    else
      FHE_UNTT_A(false, false);
  }
#undef FHE_UNTT_A
#undef FHE_UNTT
}

void launch_uneven_digits(const DeviceTables &t, const UnevenCt *cts,
                          const UnitList &list, const UnevenPtrs &x,
                          const UnevenPtrs &digits, int maxz, int cap,
                          hipStream_t stream) {
  if (list.count == 0)
    return;
  UnevenDigitsArgs a{list, cts, x, digits, maxz};
  const dim3 grid(grid_of(list.count, cap));
#define FHE_UDIG(W)                                                            \
  hipLaunchKernelGGL(uneven_digits_kernel<W>, grid, dim3(kBlock), 0, stream,  \
                     a, t)
  if (maxz <= 1)
    FHE_UDIG(1);
  else if (maxz <= 2)
    FHE_UDIG(2);
  else if (maxz <= 4)
    FHE_UDIG(4);
  else if (maxz <= 8)
    FHE_UDIG(8);
  else
    FHE_UDIG(16);
#undef FHE_UDIG
}

void launch_uneven_extend(const DeviceTables &t, const UnevenCt *cts,
                          const UnitList &list, const UnevenPtrs &digits,
                          const UnevenPtrs &ext, int maxz, int variant,
                          int cap, hipStream_t stream) {
  if (list.count == 0)
    return;
  UnevenExtArgs a{list, cts, digits, ext, maxz};
  const dim3 grid(grid_of(list.count, cap));
#define FHE_UEXT(A, W, S)                                                      \
  hipLaunchKernelGGL((uneven_extend_kernel<A, W, S>), grid, dim3(kBlock), 0,   \
                     stream, a, t)
#define FHE_UEXT_W(W)                                                          \
  do {                                                                         \
    if (variant == kExact)                                                     \
      FHE_UEXT(kExact, W, true);                                               \
    else if (variant == kLazySmall)                                            \
      FHE_UEXT(kLazySmall, W, false);                                          \
    else                                                                       \
      FHE_UEXT(kLazyLarge, W, false);                                          \
  } while (0)
  if (maxz <= 1)
    FHE_UEXT_W(1);
  else if (maxz <= 2)
    FHE_UEXT_W(2);
  else if (maxz <= 4)
    FHE_UEXT_W(4);
  else if (maxz <= 8)
    FHE_UEXT_W(8);
  else
    FHE_UEXT_W(16);
#undef FHE_UEXT_W
#undef FHE_UEXT
}

void launch_uneven_ks(const DeviceTables &t, const UnevenCt *cts,
                      const UnitList &list, const UnevenPtrs &ext,
                      const UnevenPtrs &own, const UnevenPtrs &acc,
                      const int64_t *key, int key_chain, int key_moduli_id,
                      bool lazy_accumulate, int variant, int cap,
                      hipStream_t stream) {
  if (list.count == 0)
    return;
  UnevenKsArgs a{list,      cts,           ext,
                 own,       acc,           key,
                 key_chain, key_moduli_id, lazy_accumulate ? 1 : 0};
  const dim3 grid(grid_of(list.count, cap));
  // variant 3: small primes with the 96-bit sums
  if (variant == kExact)
    hipLaunchKernelGGL((uneven_ks_kernel<kExact, false>), grid, dim3(kBlock),
                       0, stream, a, t);
  else if (variant == kLazySmall)
    hipLaunchKernelGGL((uneven_ks_kernel<kLazySmall, false>), grid,
                       dim3(kBlock), 0, stream, a, t);
  else if (variant == kLazyLarge)
    hipLaunchKernelGGL((uneven_ks_kernel<kLazyLarge, false>), grid,
                       dim3(kBlock), 0, stream, a, t);
  else
    hipLaunchKernelGGL((uneven_ks_kernel<kLazySmall, true>), grid,
                       dim3(kBlock), 0, stream, a, t);
}

void launch_uneven_moddown(const DeviceTables &t, const UnevenCt *cts,
                           const UnitList &list, const UnevenPtrs &acc,
                           const UnevenPtrs &d, int d_polys,
                           const UnevenPtrs &out, int cap,
                           hipStream_t stream) {
  if (list.count == 0)
    return;
  UnevenModDownArgs a{list, cts, acc, d, out, d_polys};
  const dim3 grid(grid_of(list.count, cap));
#define FHE_UMD(W)                                                             \
  do {                                                                         \
    if (t.alpha == W)                                                          \
      hipLaunchKernelGGL((uneven_moddown_kernel<W, true>), grid, dim3(kBlock), \
                         0, stream, a, t);                                     \
    else                                                                       \
      hipLaunchKernelGGL((uneven_moddown_kernel<W, false>), grid,             \
                         dim3(kBlock), 0, stream, a, t);                       \
  } while (0)
  if (t.alpha <= 1)
    FHE_UMD(1);
  else if (t.alpha <= 2)
    FHE_UMD(2);
  else if (t.alpha <= 4)
    FHE_UMD(4);
  else if (t.alpha <= 8)
    FHE_UMD(8);
  else
    FHE_UMD(16);
#undef FHE_UMD
}

void launch_uneven_automorphism(const DeviceTables &t, const UnevenCt *cts,
                                const UnitList &list, const UnevenPtrs &in,
                                const UnevenPtrs &out, uint32_t g, int cap,
                                hipStream_t stream) {
  if (list.count == 0)
    return;
  const uint32_t two_n = 2u << t.log_n;
  UnevenAutArgs a{list, cts, in, out, odd_inverse32(g) & (two_n - 1)};
  const dim3 grid(grid_of(list.count, cap));
  hipLaunchKernelGGL(uneven_automorphism_kernel, grid, dim3(kBlock), 0,
                     stream, a, t);
}

} // namespace fhe
