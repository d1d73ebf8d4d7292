// CKKS encoding and decoding on the device.
//
// Encoding (CKKSCore::encode_polynomial,
// src/ckks/core_encode.cpp:11-150): slot permutation, bit reversal, the
// complex FFT over 2 * slots points with the host-built twiddles
// (src/operation/fft.cpp:57-102), normalization and twisting
// (src/ckks/encode.cpp:264-275), rounding to integers (llround, or the
// reference's double_to_rns when the largest coefficient exceeds 2^59,
// src/arithmetic/arithmetic.cpp:242-269), the lift to every limb and the
// Montgomery entry with the level's scale constants.
//
// Decoding (CKKSCore::convert_plaintext_to_double and decode_,
// src/ckks/core_decode.cpp:10-124): make_unsigned_reduce, the CRT
// reconstruction of rns_to_double (src/arithmetic/arithmetic.cpp:169-197) in
// multiword integers with the reference's BigInt steps
// (src/arithmetic/big_int.cpp:176-225: Montgomery reduction by 2^62, the
// centered lift, and get_double's low-to-high sum), the skewer, the inverse
// FFT and the slot gather with the scale inverse.
//
// Bit-exactness with the reference's double arithmetic: every product and
// sum below is a separate IEEE operation in the reference's order (this file
// is compiled without floating-point contraction and uses the explicit
// round-to-nearest intrinsics), the cosine and sine values come from the
// host's libm exactly as the reference computes them, and complex products
// are (ac - bd, ad + bc).
#include "kernels.h"

namespace fhe {

namespace {

struct Cplx {
  double re, im;
};

__device__ __forceinline__ Cplx cmul(double ar, double ai, double br,
                                     double bi) {
  return {__dsub_rn(__dmul_rn(ar, br), __dmul_rn(ai, bi)),
          __dadd_rn(__dmul_rn(ar, bi), __dmul_rn(ai, br))};
}

__device__ __forceinline__ uint32_t bit_reverse(uint32_t x, int bits) {
  return __builtin_bitreverse32(x) >> (32 - bits);
}

// x [count][slots] (real) or [count][slots][2] (complex) into the
// bit-reversed complex array y [count][codec]: slot i goes to codec index
// p = pre[i] and its mirror codec - 1 - p (conjugated), as pre_permute
// (src/ckks/encode.cpp:156-185) followed by bit_reversal_permute.
// Also clears the per-ciphertext maxima of the twist kernel.
__global__ __launch_bounds__(kBlock) void encode_scatter_kernel(
    const double *x, Cplx *y, unsigned long long *max_bits, int count,
    int slots, int log_codec, const uint32_t *pre, int complex_input) {
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx < count)
    max_bits[idx] = 0;
  if (idx >= int64_t(count) * slots)
    return;
  const int64_t c = idx / slots;
  const int i = int(idx - c * slots);
  const int codec = 2 * slots;
  const double re = complex_input ? x[2 * idx] : x[idx];
  const double im = complex_input ? x[2 * idx + 1] : 0.0;
  const uint32_t p = pre[i];
  Cplx *yc = y + c * codec;
  yc[bit_reverse(p, log_codec)] = {re, im};
  yc[bit_reverse(uint32_t(codec) - 1 - p, log_codec)] = {
      re, complex_input ? -im : 0.0};
}

// One radix-2 stage of the in-place decimation-in-time FFT: step = 2^(s+1),
// twiddle index (j - offset) * (codec / step).
__global__ __launch_bounds__(kBlock) void encode_fft_stage_kernel(
    Cplx *y, int count, int log_codec, int s, const double *tw_cos,
    const double *tw_sin) {
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  const int half_codec = 1 << (log_codec - 1);
  if (idx >= int64_t(count) * half_codec)
    return;
  const int64_t c = idx >> (log_codec - 1);
  const int b = int(idx & (half_codec - 1));
  const int half = 1 << s;
  const int k = b & (half - 1);
  const int j = ((b >> s) << (s + 1)) + k;
  const int t = k << (log_codec - 1 - s);
  Cplx *yc = y + (c << log_codec);
  const Cplx left = yc[j], r = yc[j + half];
  const Cplx right = cmul(tw_cos[t], tw_sin[t], r.re, r.im);
  yc[j] = {__dadd_rn(left.re, right.re), __dadd_rn(left.im, right.im)};
  yc[j + half] = {__dsub_rn(left.re, right.re), __dsub_rn(left.im, right.im)};
}

// The same radix-2 stages fused: every butterfly is the stage kernel's (same
// twiddle index, the same rounded products and sums), only the passes over
// HBM change.  Stages 0 .. K-1 pair elements inside aligned blocks of 2^K,
// so fft_low_kernel runs them on one such block in LDS; stages K .. L-1 pair
// elements of equal index mod 2^K, so fft_high_kernel runs them on kFftCols
// such residues x 2^(L-K) rows in LDS.  Two passes instead of L.
constexpr int kFftLowBits = 11;   // 2^11 complex = 32 KiB of LDS
constexpr int kFftCols = 64;      // residues per block of the high pass
constexpr int kFftMaxHighRows = 64;

__device__ __forceinline__ void fft_butterfly(Cplx *sh, int a, int b, int t,
                                              const double *tw_cos,
                                              const double *tw_sin) {
  const Cplx left = sh[a], r = sh[b];
  const Cplx right = cmul(tw_cos[t], tw_sin[t], r.re, r.im);
  sh[a] = {__dadd_rn(left.re, right.re), __dadd_rn(left.im, right.im)};
  sh[b] = {__dsub_rn(left.re, right.re), __dsub_rn(left.im, right.im)};
}

// stages 0 .. bits-1 of one aligned block of 2^bits elements per block
__global__ __launch_bounds__(kBlock) void fft_low_kernel(
    Cplx *y, int log_codec, int bits, const double *tw_cos,
    const double *tw_sin) {
  __shared__ Cplx sh[1 << kFftLowBits];
  const int size = 1 << bits;
  Cplx *blk = y + (int64_t(blockIdx.x) << bits);
  for (int e = threadIdx.x; e < size; e += kBlock)
    sh[e] = blk[e];
  __syncthreads();
  for (int s = 0; s < bits; s++) {
    const int half = 1 << s;
    for (int bf = threadIdx.x; bf < size / 2; bf += kBlock) {
      // the block starts at a multiple of 2^bits: the index inside the
      // block has the message index's low bits
      const int k = bf & (half - 1);
      const int j = ((bf >> s) << (s + 1)) + k;
      fft_butterfly(sh, j, j + half, k << (log_codec - 1 - s), tw_cos,
                    tw_sin);
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < size; e += kBlock)
    blk[e] = sh[e];
}

// stages low_bits .. log_codec-1 of kFftCols consecutive residues mod
// 2^low_bits (all their 2^(log_codec - low_bits) rows) per block
__global__ __launch_bounds__(kBlock) void fft_high_kernel(
    Cplx *y, int log_codec, int low_bits, const double *tw_cos,
    const double *tw_sin) {
  __shared__ Cplx sh[kFftMaxHighRows * kFftCols];
  const int rows_log = log_codec - low_bits;
  const int rows = 1 << rows_log;
  const int64_t blocks_per_msg = (int64_t(1) << low_bits) / kFftCols;
  const int64_t c = blockIdx.x / blocks_per_msg;
  const int r0 = int(blockIdx.x % blocks_per_msg) * kFftCols;
  Cplx *msg = y + (c << log_codec);
  for (int e = threadIdx.x; e < rows * kFftCols; e += kBlock) {
    const int row = e / kFftCols, col = e % kFftCols;
    sh[e] = msg[(int64_t(row) << low_bits) + r0 + col];
  }
  __syncthreads();
  for (int s = low_bits; s < log_codec; s++) {
    const int rh = 1 << (s - low_bits); // rows between the pair
    for (int bf = threadIdx.x; bf < (rows / 2) * kFftCols; bf += kBlock) {
      const int col = bf % kFftCols, rb = bf / kFftCols;
      const int kr = rb & (rh - 1);
      const int jr = ((rb >> (s - low_bits)) << (s - low_bits + 1)) + kr;
      // j mod 2^s = kr * 2^low_bits + (r0 + col)
      const int k = (kr << low_bits) + r0 + col;
      fft_butterfly(sh, jr * kFftCols + col, (jr + rh) * kFftCols + col,
                    k << (log_codec - 1 - s), tw_cos, tw_sin);
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < rows * kFftCols; e += kBlock) {
    const int row = e / kFftCols, col = e % kFftCols;
    msg[(int64_t(row) << low_bits) + r0 + col] = sh[e];
  }
}

// normalize (src/operation/fft.cpp:50-55) and apply_twister
// (src/ckks/encode.cpp:264-275): out[i] = (re_i / codec * cos_i - im_i /
// codec * sin_i) * scale; the per-ciphertext max |out| as ordered bits.
__global__ __launch_bounds__(kBlock) void encode_twist_kernel(
    const Cplx *y, double *out, unsigned long long *max_bits, int count,
    int log_codec, const double *cos_t, const double *sin_t, double scale) {
  // codec >= kBlock, so a block's elements belong to one message
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  const int codec = 1 << log_codec;
  const int i = int(idx & (codec - 1));
  const double size = double(codec);
  const double re = __ddiv_rn(y[idx].re, size);
  const double im = __ddiv_rn(y[idx].im, size);
  const double v = __dmul_rn(
      __dsub_rn(__dmul_rn(re, cos_t[i]), __dmul_rn(im, sin_t[i])), scale);
  out[idx] = v;
  // |v| compares as its bit pattern: wave maximum by butterfly shuffles,
  // then one atomic per wave
  unsigned long long m = (unsigned long long)__double_as_longlong(fabs(v));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(m, off, 64);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0)
    atomicMax(max_bits + (idx >> log_codec), m);
}

// All log_codec stages on `count` messages [count][2^log_codec] in place:
// the fused passes when their shapes fit, the per-stage kernel otherwise.
void launch_fft(Cplx *y, int count, int log_codec, const double *tw_cos,
                const double *tw_sin, hipStream_t stream) {
  const int low = log_codec < kFftLowBits ? log_codec : kFftLowBits;
  const int high_rows = 1 << (log_codec - low);
  const bool fused = log_codec >= 8 && high_rows <= kFftMaxHighRows &&
                     (log_codec == low || (1 << low) % kFftCols == 0);
  if (!fused) {
    const int64_t half_codec = int64_t(1) << (log_codec - 1);
    for (int s = 0; s < log_codec; s++)
      hipLaunchKernelGGL(encode_fft_stage_kernel,
                         dim3(grid_for(int64_t(count) * half_codec)),
                         dim3(kBlock), 0, stream, y, count, log_codec, s,
                         tw_cos, tw_sin);
    return;
  }
  hipLaunchKernelGGL(fft_low_kernel,
                     dim3(unsigned(int64_t(count) << (log_codec - low))),
                     dim3(kBlock), 0, stream, y, log_codec, low, tw_cos,
                     tw_sin);
  if (log_codec > low)
    hipLaunchKernelGGL(fft_high_kernel,
                       dim3(unsigned((int64_t(count) << low) / kFftCols)),
                       dim3(kBlock), 0, stream, y, log_codec, low, tw_cos,
                       tw_sin);
}

struct ConvertArgs {
  const double *v;            // [count][codec]
  const unsigned long long *max_bits;
  int64_t *out;               // [count][out_limbs][N]
  int count, log_n, out_limbs;
  int first;                  // limbs before the encoded ones (zero-filled)
  int moduli_id;              // prime of encoded limb 0
  int64_t c[kMaxEwLimbs];     // r_mult_scale_after_tile per encoded limb
};

// positive_modulo (src/arithmetic/arithmetic.cpp:271-273)
__device__ __forceinline__ int64_t positive_mod(int64_t x, int64_t q) {
  return (q + x % q) % q;
}

__global__ __launch_bounds__(kBlock) void encode_convert_kernel(
    ConvertArgs a, DeviceTables t) {
  const int64_t n = int64_t(1) << a.log_n;
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= int64_t(a.count) * a.out_limbs * n)
    return;
  const int64_t coef = idx & (n - 1);
  const int64_t cl = idx >> a.log_n;
  const int l = int(cl % a.out_limbs);
  const int64_t c = cl / a.out_limbs;
  if (l < a.first) {
    a.out[idx] = 0;
    return;
  }
  const int e = l - a.first;
  const int p = a.moduli_id + e;
  const int64_t q = int64_t(t.q[p]);
  const double x = a.v[(c << a.log_n) + coef];
  const double max_abs = __longlong_as_double((long long)a.max_bits[c]);
  int64_t y;
  if (max_abs > double(uint64_t(1) << 59)) {
    // double_to_rns
    if (fabs(x) < double(int64_t(1) << 62)) {
      y = positive_mod(int64_t(__dadd_rn(x, 0.5)), q);
    } else {
      int exponent;
      const double mantissa = frexp(__dadd_rn(x, 0.5), &exponent);
      int64_t r = positive_mod(
          int64_t(__dmul_rn(mantissa, double(int64_t(1) << 52))), q);
      for (int le = 52; le < exponent; le++) {
        r <<= 1;
        r %= q;
      }
      y = r;
    }
  } else {
    // llround, then tile_unsigned_reduce_2q_to_q
    const int64_t r = llround(x);
    y = r < 0 ? r + q : guard(r, q);
  }
  a.out[idx] = mont_mult(y, a.c[e], t.q[p], t.k[p]);
}

} // namespace

void launch_encode(const DeviceTables &t, const EncodeDevice &e,
                   const double *x, int complex_input, int count,
                   double scale_before_tile, int out_limbs, int first,
                   int moduli_id, const int64_t *r_mult, int64_t *out,
                   void *work, hipStream_t stream) {
  const int codec = 1 << e.log_codec;
  const int slots = codec / 2;
  Cplx *y = static_cast<Cplx *>(work);
  double *v = reinterpret_cast<double *>(y + int64_t(count) * codec);
  auto *max_bits =
      reinterpret_cast<unsigned long long *>(v + int64_t(count) * codec);
  hipLaunchKernelGGL(encode_scatter_kernel,
                     dim3(grid_for(int64_t(count) * slots)), dim3(kBlock), 0,
                     stream, x, y, max_bits, count, slots, e.log_codec, e.pre,
                     complex_input);
  launch_fft(y, count, e.log_codec, e.fft_cos, e.fft_sin, stream);
  // the twist kernel's grid covers count * codec exactly (codec is a
  // multiple of kBlock): every thread of a wave takes part in its shuffles
  hipLaunchKernelGGL(encode_twist_kernel,
                     dim3(unsigned((int64_t(count) * codec) / kBlock)),
                     dim3(kBlock), 0,
                     stream, y, v, max_bits, count, e.log_codec, e.twist_cos,
                     e.twist_sin, scale_before_tile);
  ConvertArgs a{};
  a.v = v;
  a.max_bits = max_bits;
  a.out = out;
  a.count = count;
  a.log_n = t.log_n;
  a.out_limbs = out_limbs;
  a.first = first;
  a.moduli_id = moduli_id;
  for (int l = 0; l < out_limbs - first; l++)
    a.c[l] = r_mult[l];
  hipLaunchKernelGGL(encode_convert_kernel,
                     dim3(grid_for((int64_t(count) * out_limbs) << t.log_n)),
                     dim3(kBlock), 0, stream, a, t);
}

size_t encode_work_bytes(const EncodeDevice &e, int count) {
  const int64_t codec = int64_t(1) << e.log_codec;
  return size_t(count) * codec * (16 + 8) + size_t(count) * 8 + 256;
}

namespace {

// One coefficient of every plaintext: the signed integer the limbs encode
// (CRT, centered), as a double.
template <int MAXW>
__global__ __launch_bounds__(kBlock) void decode_crt_kernel(
    const int64_t *in, double *v, int count, int limbs, int mid,
    DecodeCrtDevice crt, DeviceTables t) {
  const int64_t n = int64_t(1) << t.log_n;
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= int64_t(count) * n)
    return;
  const int64_t c = idx >> t.log_n;
  const int64_t coef = idx & (n - 1);
  const int64_t *x = in + c * limbs * n + coef;
  if (limbs == 1) {
    // make_signed (src/montgomery/unit.cpp:36-42; unsigned compare) and the
    // int64 -> double cast
    const int64_t q = int64_t(t.q[mid]);
    const int64_t xv = x[0];
    const int64_t y = uint64_t(xv) <= t.qh[mid] ? xv : xv - q;
    v[idx] = __ll2double_rn(y);
    return;
  }
  const int words = crt.words;
  uint64_t acc[MAXW + 2];
#pragma unroll
  for (int w = 0; w < MAXW + 2; w++)
    acc[w] = 0;
  for (int l = 0; l < limbs; l++) {
    const int p = mid + l;
    const int64_t q = int64_t(t.q[p]);
    int64_t xl = x[int64_t(l) * n];
    xl = xl < 0 ? xl + q : guard(xl, q);
    // x 2^62 mod q, canonical: a Montgomery product with R^2 mod q
    const uint64_t xr =
        uint64_t(guard(mont_mult(xl, t.r2[p], t.q[p], t.k[p]), q));
    const uint64_t *rv = crt.reverser + int64_t(l) * words;
    uint64_t carry = 0;
#pragma unroll
    for (int w = 0; w < MAXW + 2; w++) {
      if (w < words) {
        const uint64_t r = rv[w];
        const uint64_t lo = r * xr;
        uint64_t hi = mulhi64(r, xr);
        uint64_t s = lo + acc[w];
        hi += s < lo;
        const uint64_t s2 = s + carry;
        hi += s2 < s;
        acc[w] = s2;
        carry = hi;
      } else {
        const uint64_t s = acc[w] + carry;
        carry = s < carry;
        acc[w] = s;
      }
    }
  }
  const uint64_t orig0 = acc[0];
  // subtract U = Q 2^62 while acc >= U (acc < limbs U)
  for (int it = 0; it <= limbs; it++) {
    bool geq = true, decided = false;
#pragma unroll
    for (int w = MAXW + 1; w >= 0; w--) {
      const uint64_t u = w <= words ? crt.upper[w] : 0;
      if (!decided && acc[w] != u) {
        geq = acc[w] > u;
        decided = true;
      }
    }
    if (!geq)
      break;
    uint64_t borrow = 0;
#pragma unroll
    for (int w = 0; w < MAXW + 2; w++) {
      const uint64_t u = w <= words ? crt.upper[w] : 0;
      const uint64_t d = acc[w] - u;
      const uint64_t d2 = d - borrow;
      borrow = (acc[w] < u) | (d < borrow);
      acc[w] = d2;
    }
  }
  // + Q m with m = (acc_0 k) mod 2^62, then the exact shift by 62
  const uint64_t m = (orig0 * crt.k) & ((uint64_t(1) << 62) - 1);
  {
    uint64_t carry = 0;
#pragma unroll
    for (int w = 0; w < MAXW + 2; w++) {
      if (w < words) {
        const uint64_t qw = crt.q_product[w];
        const uint64_t lo = qw * m;
        uint64_t hi = mulhi64(qw, m);
        const uint64_t s = lo + acc[w];
        hi += s < lo;
        const uint64_t s2 = s + carry;
        hi += s2 < s;
        acc[w] = s2;
        carry = hi;
      } else {
        const uint64_t s = acc[w] + carry;
        carry = s < carry;
        acc[w] = s;
      }
    }
  }
  uint64_t quo[MAXW + 1];
#pragma unroll
  for (int w = 0; w < MAXW + 1; w++)
    quo[w] = (acc[w] >> 62) | (acc[w + 1] << 2);
  // quo < 2Q: one conditional subtraction of Q, then the centered lift
  auto compare = [&](const uint64_t *b) {
    // quo >= b (b has `words` words)
    bool geq = true, decided = false;
#pragma unroll
    for (int w = MAXW; w >= 0; w--) {
      const uint64_t bw = w < words ? b[w] : 0;
      if (!decided && quo[w] != bw) {
        geq = quo[w] > bw;
        decided = true;
      }
    }
    return geq;
  };
  if (compare(crt.q_product)) {
    uint64_t borrow = 0;
#pragma unroll
    for (int w = 0; w < MAXW + 1; w++) {
      const uint64_t bw = w < words ? crt.q_product[w] : 0;
      const uint64_t d = quo[w] - bw;
      const uint64_t d2 = d - borrow;
      borrow = (quo[w] < bw) | (d < borrow);
      quo[w] = d2;
    }
  }
  const bool negative = compare(crt.q_half);
  if (negative) {
    // Q - quo
    uint64_t borrow = 0;
#pragma unroll
    for (int w = 0; w < MAXW + 1; w++) {
      const uint64_t bw = w < words ? crt.q_product[w] : 0;
      const uint64_t d = bw - quo[w];
      const uint64_t d2 = d - borrow;
      borrow = (bw < quo[w]) | (d < borrow);
      quo[w] = d2;
    }
  }
  // get_double over the words up to the highest nonzero one:
  // y += double(word_i) * 2^(64 i), low to high
  int size = 1;
#pragma unroll
  for (int w = 1; w < MAXW + 1; w++)
    if (quo[w] != 0)
      size = w + 1;
  double y = 0.0;
#pragma unroll
  for (int w = 0; w < MAXW + 1; w++) {
    if (w < size) {
      const double pw = 64 * w < 1024 ? ldexp(1.0, 64 * w) : __longlong_as_double(
                                                               0x7ff0000000000000ll);
      y = __dadd_rn(y, __dmul_rn(__ull2double_rn(quo[w]), pw));
    }
  }
  v[idx] = negative ? -y : y;
}

// skewer (src/ckks/encode.cpp:289-297) into bit-reversed complex order
__global__ __launch_bounds__(kBlock) void decode_skew_kernel(
    const double *v, Cplx *y, int count, int log_codec, const double *cos_t,
    const double *sin_t) {
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= (int64_t(count) << log_codec))
    return;
  const int64_t c = idx >> log_codec;
  const int i = int(idx & ((int64_t(1) << log_codec) - 1));
  const double x = v[idx];
  y[(c << log_codec) + bit_reverse(uint32_t(i), log_codec)] = {
      __dmul_rn(x, cos_t[i]), __dmul_rn(x, sin_t[i])};
}

// permute_after_ifft (src/ckks/encode.cpp:481-503)
__global__ __launch_bounds__(kBlock) void decode_gather_kernel(
    const Cplx *y, double *out, int count, int log_codec, const uint32_t *post,
    double scale_inverse, int complex_output) {
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= (int64_t(count) << log_codec))
    return;
  const int64_t c = idx >> log_codec;
  const int i = int(idx & ((int64_t(1) << log_codec) - 1));
  const int slots = 1 << (log_codec - 1);
  const uint32_t p = post[i];
  if (p >= uint32_t(slots))
    return;
  const int64_t o = c * slots + p;
  if (complex_output) {
    out[2 * o] = __dmul_rn(y[idx].re, scale_inverse);
    out[2 * o + 1] = __dmul_rn(y[idx].im, scale_inverse);
  } else {
    out[o] = __dmul_rn(y[idx].re, scale_inverse);
  }
}

} // namespace

size_t decode_work_bytes(const DecodeDevice &d, int count) {
  const int64_t codec = int64_t(1) << d.log_codec;
  return size_t(count) * codec * (16 + 8) + 256;
}

bool decode_crt_supported(int words) { return words <= kMaxDecodeWords; }

void launch_decode(const DeviceTables &t, const DecodeDevice &d,
                   const DecodeCrtDevice &crt, const int64_t *in, int count,
                   int limbs, int moduli_id, double scale_inverse,
                   int complex_output, double *out, void *work,
                   hipStream_t stream) {
  const int codec = 1 << d.log_codec;
  double *v = static_cast<double *>(work);
  Cplx *y = reinterpret_cast<Cplx *>(v + int64_t(count) * codec);
  const dim3 grid(grid_for(int64_t(count) * codec));
  if (crt.words <= 8)
    hipLaunchKernelGGL(decode_crt_kernel<8>, grid, dim3(kBlock), 0, stream, in,
                       v, count, limbs, moduli_id, crt, t);
  else if (crt.words <= 16)
    hipLaunchKernelGGL(decode_crt_kernel<16>, grid, dim3(kBlock), 0, stream,
                       in, v, count, limbs, moduli_id, crt, t);
  else if (crt.words <= 24)
    hipLaunchKernelGGL(decode_crt_kernel<24>, grid, dim3(kBlock), 0, stream,
                       in, v, count, limbs, moduli_id, crt, t);
  else
    hipLaunchKernelGGL(decode_crt_kernel<kMaxDecodeWords>, grid, dim3(kBlock),
                       0, stream, in, v, count, limbs, moduli_id, crt, t);
  hipLaunchKernelGGL(decode_skew_kernel, grid, dim3(kBlock), 0, stream, v, y,
                     count, d.log_codec, d.skew_cos, d.skew_sin);
  launch_fft(y, count, d.log_codec, d.ifft_cos, d.ifft_sin, stream);
  hipLaunchKernelGGL(decode_gather_kernel, grid, dim3(kBlock), 0, stream, y,
                     out, count, d.log_codec, d.post, scale_inverse,
                     complex_output);
}

} // namespace fhe
