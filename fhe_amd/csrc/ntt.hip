// Batched negacyclic NTT / INTT for gfx950.
//
// Butterflies are exactly the reference's radix-2 formulas
// (src/operation/ntt.cpp:46-134 Cooley-Tukey, :177-272 Gentleman-Sande), so
// every lazy representative in [0, 2q) matches.  The transform of one limb is
// split into at most two passes over HBM: each pass owns a contiguous range of
// index bits, loads a 4096-element tile of independent sub-transforms into
// LDS and runs all of that range's radix-2 stages there.
//
//   forward  stage s flips index bit log_n-1-s; group of element i uses
//            root[2^s - 1 + (i >> (log_n - s))]
//   inverse  stage t flips bit t; group uses
//            root_inverse[N - N/2^t + (i >> (t + 1))]
//
// A pass covering bits [b, b+L) sees sub-transform elements
//   i = (hi << (b+L)) | (j << b) | lo,   j in [0, 2^L)
// and in both directions the twiddle offset is i >> (b + u + 1) for local bit u.
#include <stdexcept>

#include "kernels.h"
#include "ntt_r16.h"

namespace fhe {

namespace {

constexpr int kTile = 4096; // elements per workgroup (32 KiB of LDS)

struct PassArgs {
  const int64_t *src;
  int64_t *dst;
  int64_t total_subs; // sub-transforms over the whole batch
  int limbs;          // limbs per polynomial: limb f uses prime mid + f % limbs
  int moduli_id;
  int low_bit;        // b
  int bits;           // L <= 12
  const int64_t *pre; // forward first pass: multiply by pre[prime] on load
  int post;           // inverse last pass: InttPost
  int last;           // the pass completes the transform
  int lazy_period;    // lazy arithmetic for polynomial index % period <
  int lazy_count;     //   count (radix-16 passes only; 0: all exact)
  int lazy_part = 0;  // LazySpec::part
};

template <bool INV>
__global__ __launch_bounds__(kBlock) void ntt_pass_kernel(PassArgs a,
                                                          DeviceTables t) {
  __shared__ int64_t lds[kTile];
  const int log_n = t.log_n;
  const int64_t n = int64_t(1) << log_n;
  const int L = a.bits, b = a.low_bit;
  const int sub = 1 << L;
  const int C = kTile >> L;              // sub-transforms per tile
  const int sub_log = log_n - L;         // log2(sub-transforms per limb)
  const int64_t tile0 = int64_t(blockIdx.x) * C;
  const bool columns = b > 0;            // strided sub-transforms
  const int tid = threadIdx.x;

  // sub-transform sp, element j -> limb-local index i, flat limb f
  auto locate = [&](int64_t sp, int j, int64_t &f) -> int64_t {
    f = sp >> sub_log;
    const int64_t r = sp & ((int64_t(1) << sub_log) - 1);
    const int64_t lo = r & ((int64_t(1) << b) - 1);
    const int64_t hi = r >> b;
    return (hi << (b + L)) | (int64_t(j) << b) | lo;
  };

  for (int e = tid; e < kTile; e += kBlock) {
    const int c = columns ? e % C : e / sub;
    const int j = columns ? e / C : e % sub;
    const int64_t sp = tile0 + c;
    int64_t v = 0;
    if (sp < a.total_subs) {
      int64_t f;
      const int64_t i = locate(sp, j, f);
      v = a.src[f * n + i];
      if (a.pre) {
        const int p = a.moduli_id + int(f % a.limbs);
        v = mont_mult_cs(a.pre[p], v, t.q[p], t.k[p]);
      }
    }
    lds[j * C + c] = v;
  }
  __syncthreads();

  for (int step = 0; step < L; step++) {
    const int u = INV ? step : L - 1 - step;
    for (int bf = tid; bf < kTile / 2; bf += kBlock) {
      const int c = bf % C;
      const int kk = bf / C;
      const int64_t sp = tile0 + c;
      if (sp >= a.total_subs)
        continue;
      const int jl = ((kk >> u) << (u + 1)) | (kk & ((1 << u) - 1));
      const int jh = jl | (1 << u);
      int64_t f;
      const int64_t i = locate(sp, jl, f);
      const int p = a.moduli_id + int(f % a.limbs);
      const int64_t q = int64_t(t.q[p]);
      const int64_t q2 = int64_t(t.q2[p]);
      const uint64_t k = t.k[p];
      const int64_t x = lds[jl * C + c];
      const int64_t y = lds[jh * C + c];
      const int64_t off = i >> (b + u + 1);
      if (!INV) {
        const int s = log_n - 1 - (b + u);
        const int64_t w = t.root[int64_t(p) * n + (int64_t(1) << s) - 1 + off];
        int64_t xa = x, ya = y;
        ct_butterfly(xa, ya, w, q2, q, k);
        lds[jl * C + c] = xa;
        lds[jh * C + c] = ya;
      } else {
        const int tt = b + u;
        const int64_t w =
            t.root_inverse[int64_t(p) * n + n - (n >> tt) + off];
        int64_t xa = x, ya = y;
        gs_butterfly(xa, ya, w, q2, q, k);
        lds[jl * C + c] = xa;
        lds[jh * C + c] = ya;
      }
    }
    __syncthreads();
  }

  for (int e = tid; e < kTile; e += kBlock) {
    const int c = columns ? e % C : e / sub;
    const int j = columns ? e / C : e % sub;
    const int64_t sp = tile0 + c;
    if (sp >= a.total_subs)
      continue;
    int64_t f;
    const int64_t i = locate(sp, j, f);
    int64_t v = lds[j * C + c];
    if (INV && a.post != kInttRaw) {
      const int p = a.moduli_id + int(f % a.limbs);
      const uint64_t q = t.q[p], k = t.k[p];
      v = mont_exit(mont_mult_cs(t.ninv[p], v, q, k), q, k);
      if (a.post == kInttExitReduce)
        v = guard(v, int64_t(q));
    }
    a.dst[f * n + i] = v;
  }
}

// Register-resident 8-bit pass (r16_limb_pass, ntt_r16.h) on limb l of
// polynomial `poly` of the batch.
template <bool INV, bool COLS, int A>
__device__ __forceinline__ void r16_pass(const PassArgs &a,
                                         const DeviceTables &t, int64_t *lds,
                                         int64_t poly, int tile, int l) {
  const int64_t limb = (poly * a.limbs + l) << t.log_n;
  r16_limb_pass<INV, COLS, A>(a.src + limb, a.dst + limb, a.moduli_id + l,
                              tile, a.low_bit, a.pre, a.post, a.last != 0, t,
                              lds);
}

// Register-resident pass kernel with one arithmetic A per launch (a single
// code path per kernel keeps the instruction footprint and the register
// allocation to that path).  The launch covers the polynomials
// poly = (j / sel_count) * sel_period + sel_offset + j % sel_count,
// j < sel_polys, and the limbs l0 .. l0 + limbs_in_launch - 1.
struct PolySelect {
  int polys;   // selected polynomials
  int period, count, offset;
  int l0;      // first limb of the launch
};

template <bool INV, bool COLS, int A>
__global__ __launch_bounds__(kBlock) void ntt_radix16_kernel(
    PassArgs a, DeviceTables t, PolySelect sel, int tiles_per_limb,
    int swizzle) {
  __shared__ int64_t lds[kR16Lds];
  // block -> (limb l, tile, polynomial), polynomial fastest so blocks sharing
  // a prime's twiddles run together; with swizzle each of the 8 XCDs (blocks
  // are dealt round-robin to XCDs) gets one contiguous range of that order.
  int64_t w = blockIdx.x;
  if (swizzle) {
    const int64_t per = int64_t(gridDim.x) >> 3;
    w = int64_t(blockIdx.x & 7) * per + (blockIdx.x >> 3);
  }
  const int64_t j = w % sel.polys;
  const int64_t rest = w / sel.polys;
  const int tile = int(rest % tiles_per_limb);
  const int l = sel.l0 + int(rest / tiles_per_limb);
  const int64_t poly =
      (j / sel.count) * sel.period + sel.offset + (j % sel.count);
  r16_pass<INV, COLS, A>(a, t, lds, poly, tile, l);
}

// The same pass with the arithmetic chosen per block: polynomials whose index
// modulo a.lazy_period is below a.lazy_count take the lazy arithmetic of
// their prime's size class, the others the exact one (one launch for all).
template <bool INV, bool COLS>
__global__ __launch_bounds__(kBlock) void ntt_radix16_mixed_kernel(
    PassArgs a, DeviceTables t, int polys, int tiles_per_limb, int swizzle) {
  __shared__ int64_t lds[kR16Lds];
  int64_t w = blockIdx.x;
  if (swizzle) {
    const int64_t per = int64_t(gridDim.x) >> 3;
    w = int64_t(blockIdx.x & 7) * per + (blockIdx.x >> 3);
  }
  const int64_t poly = w % polys;
  const int64_t rest = w / polys;
  const int tile = int(rest % tiles_per_limb);
  const int l = int(rest / tiles_per_limb);
  const bool lazy = a.lazy_period > 0 && (poly % a.lazy_period) < a.lazy_count;
  if (!lazy)
    r16_pass<INV, COLS, kExact>(a, t, lds, poly, tile, l);
  else if (t.q[a.moduli_id + l] < kLazySmallBound)
    r16_pass<INV, COLS, kLazySmall>(a, t, lds, poly, tile, l);
  else
    r16_pass<INV, COLS, kLazyLarge>(a, t, lds, poly, tile, l);
}

// Split log_n into the bit ranges of at most two passes (each <= 12 bits).
void split_bits(int log_n, int &first, int &second) {
  if (log_n <= 12) {
    first = log_n;
    second = 0;
  } else {
    first = log_n / 2;
    second = log_n - first;
  }
}

void launch_pass(bool inverse, const PassArgs &a, const DeviceTables &t,
                 hipStream_t stream) {
  if (a.bits == 8 && t.log_n >= 12) {
    const int64_t total_limbs = a.total_subs >> (t.log_n - 8);
    const int polys = int(total_limbs / a.limbs);
    const int tiles = int(int64_t(1) << (t.log_n - 12));
    const bool cols = a.low_bit > 0;
    auto launch = [&](int arith, PolySelect sel, int limbs_in_launch) {
      const int64_t blocks = int64_t(sel.polys) * tiles * limbs_in_launch;
      if (blocks == 0)
        return;
      const int swizzle = (blocks % 8 == 0) ? 1 : 0;
#define FHE_R16(I, C, A)                                                       \
  hipLaunchKernelGGL((ntt_radix16_kernel<I, C, A>), dim3(unsigned(blocks)),   \
                     dim3(kBlock), 0, stream, a, t, sel, tiles, swizzle)
#define FHE_R16_A(I, C)                                                        \
  do {                                                                         \
    if (arith == kExact)                                                       \
      FHE_R16(I, C, kExact);                                                   \
    else if (arith == kLazySmall)                                              \
      FHE_R16(I, C, kLazySmall);                                               \
    else                                                                       \
      FHE_R16(I, C, kLazyLarge);                                               \
  } while (0)
      if (inverse) {
        if (cols)
          FHE_R16_A(true, true);
        else
          FHE_R16_A(true, false);
      } else {
        if (cols)
          FHE_R16_A(false, true);
        else
          FHE_R16_A(false, false);
      }
#undef FHE_R16_A
#undef FHE_R16
    };
    if (!t.split_variants) {
      const int64_t blocks = total_limbs * tiles;
      const int swizzle = (blocks % 8 == 0) ? 1 : 0;
#define FHE_R16M(I, C)                                                         \
  hipLaunchKernelGGL((ntt_radix16_mixed_kernel<I, C>), dim3(unsigned(blocks)),\
                     dim3(kBlock), 0, stream, a, t, polys, tiles, swizzle)
      if (inverse) {
        if (cols)
          FHE_R16M(true, true);
        else
          FHE_R16M(true, false);
      } else {
        if (cols)
          FHE_R16M(false, true);
        else
          FHE_R16M(false, false);
      }
#undef FHE_R16M
      return;
    }
    const int period = a.lazy_period > 0 ? a.lazy_period : 1;
    const int lazy_count = a.lazy_period > 0 ? a.lazy_count : 0;
    // polynomials in whole periods (every caller passes whole periods)
    if (polys % period != 0)
      throw std::invalid_argument(
          "NTT: the polynomial count is not a whole number of lazy periods");
    const int periods = polys / period;
    // exact polynomials: all limbs in one launch
    if (lazy_count < period && a.lazy_part != 2)
      launch(kExact,
             PolySelect{periods * (period - lazy_count), period,
                        period - lazy_count, lazy_count, 0},
             a.limbs);
    // lazy polynomials: one launch per run of limbs of one prime class
    if (lazy_count > 0 && a.lazy_part != 1) {
      int l = 0;
      while (l < a.limbs) {
        const bool small = t_host_q_small(t, a.moduli_id + l);
        int e = l + 1;
        while (e < a.limbs && t_host_q_small(t, a.moduli_id + e) == small)
          e++;
        launch(small ? kLazySmall : kLazyLarge,
               PolySelect{periods * lazy_count, period, lazy_count, 0, l},
               e - l);
        l = e;
      }
    }
    return;
  }
  const int C = kTile >> a.bits;
  const unsigned blocks = unsigned((a.total_subs + C - 1) / C);
  if (inverse)
    hipLaunchKernelGGL(ntt_pass_kernel<true>, dim3(blocks), dim3(kBlock), 0,
                       stream, a, t);
  else
    hipLaunchKernelGGL(ntt_pass_kernel<false>, dim3(blocks), dim3(kBlock), 0,
                       stream, a, t);
}

} // namespace

bool ntt_parts_supported(const DeviceTables &t) {
  int l1, l2;
  split_bits(t.log_n, l1, l2);
  return t.split_variants && l1 == 8 && l2 == 8;
}

void ntt_forward(const DeviceTables &t, const int64_t *src, int64_t *dst,
                 int64_t total_limbs, int limbs, int moduli_id,
                 const int64_t *pre, hipStream_t stream, LazySpec lazy) {
  int l1, l2;
  split_bits(t.log_n, l1, l2);
  // first pass: the high bits (stages 0 .. l1-1)
  PassArgs a{src,     dst,      total_limbs << (t.log_n - l1),
             limbs,   moduli_id, t.log_n - l1,
             l1,      pre,       kInttRaw,
             l2 == 0, lazy.period, lazy.count, lazy.part};
  launch_pass(false, a, t, stream);
  if (l2 > 0) {
    PassArgs b2{dst,   dst,      total_limbs << (t.log_n - l2),
                limbs, moduli_id, 0,
                l2,    nullptr,   kInttRaw,
                1,     lazy.period, lazy.count, lazy.part};
    launch_pass(false, b2, t, stream);
  }
}

void ntt_inverse(const DeviceTables &t, const int64_t *src, int64_t *dst,
                 int64_t total_limbs, int limbs, int moduli_id, int post,
                 hipStream_t stream, LazySpec lazy) {
  int l1, l2;
  split_bits(t.log_n, l1, l2);
  // first pass: the low bits (stages with span 1 .. 2^(l1-1))
  PassArgs a{src,     dst,      total_limbs << (t.log_n - l1),
             limbs,   moduli_id, 0,
             l1,      nullptr,   l2 > 0 ? int(kInttRaw) : post,
             l2 == 0, lazy.period, lazy.count, lazy.part};
  launch_pass(true, a, t, stream);
  if (l2 > 0) {
    PassArgs b2{dst,   dst,      total_limbs << (t.log_n - l2),
                limbs, moduli_id, l1,
                l2,    nullptr,   post,
                1,     lazy.period, lazy.count, lazy.part};
    launch_pass(true, b2, t, stream);
  }
}

void ntt_forward_first(const DeviceTables &t, const int64_t *src, int64_t *dst,
                       int64_t total_limbs, int limbs, int moduli_id,
                       const int64_t *pre, hipStream_t stream,
                       LazySpec lazy) {
  int l1, l2;
  split_bits(t.log_n, l1, l2);
  PassArgs a{src,     dst,      total_limbs << (t.log_n - l1),
             limbs,   moduli_id, t.log_n - l1,
             l1,      pre,       kInttRaw,
             l2 == 0, lazy.period, lazy.count, lazy.part};
  launch_pass(false, a, t, stream);
}

void ntt_forward_second(const DeviceTables &t, const int64_t *src,
                        int64_t *dst, int64_t total_limbs, int limbs,
                        int moduli_id, hipStream_t stream, LazySpec lazy) {
  int l1, l2;
  split_bits(t.log_n, l1, l2);
  if (l2 == 0) {
    // single-pass sizes: the first pass was the whole transform
    if (src != dst)
      (void)hipMemcpyAsync(dst, src, size_t(total_limbs << t.log_n) * 8,
                           hipMemcpyDeviceToDevice, stream);
    return;
  }
  PassArgs b2{src,   dst,      total_limbs << (t.log_n - l2),
              limbs, moduli_id, 0,
              l2,    nullptr,   kInttRaw,
              1,     lazy.period, lazy.count, lazy.part};
  launch_pass(false, b2, t, stream);
}

void ntt_inverse_second(const DeviceTables &t, const int64_t *src,
                        int64_t *dst, int64_t total_limbs, int limbs,
                        int moduli_id, int post, hipStream_t stream,
                        LazySpec lazy) {
  int l1, l2;
  split_bits(t.log_n, l1, l2);
  PassArgs a{src,   dst,      total_limbs << (t.log_n - l2),
             limbs, moduli_id, l1,
             l2,    nullptr,   post,
             1,     lazy.period, lazy.count, lazy.part};
  launch_pass(true, a, t, stream);
}

} // namespace fhe
