// Tensor product fused with the transforms around it (N = 2^16): the second
// (low-bit) pass of the forward NTT of both input ciphertexts, the tensor
// z0 = a0 b0, z1 = reduce(a0 b1 + a1 b0), z2 = a1 b1
// (CKKSCore::multiply, src/ckks/core.cpp:284-314) and the first (low-bit)
// pass of the inverse NTT of the three products, in registers.  A block owns
// one 4096-coefficient tile of one limb of one ciphertext pair; the NTT
// values of the four inputs exist only in registers and the trio is written
// once, after its first inverse pass.
//
// a1, b1 and z2 use the exact arithmetic: z2's inverse transform feeds the
// representative-sensitive ModUp.  a0, b0, z0 and z1 may use the lazy
// arithmetic, because z0 and z1 are canonicalized by the relinearization add.
#include "kernels.h"
#include "ntt_r16.h"

namespace fhe {

namespace {

struct TensorArgs {
  const int64_t *a; // [count][2][limbs][N] after the forward high-bit pass
  const int64_t *b;
  int64_t *trio;    // [count][3][limbs][N] after the inverse low-bit pass
  int count, limbs, moduli_id, lazy, swizzle;
};

// Forward low-bit pass of one input tile: load ROWS, bits 7..4, exchange to
// COLS, bits 3..0; the thread then holds coefficients 16 r + m of column c.
template <int A>
__device__ __forceinline__ void forward_tile(const int64_t *src,
                                             int64_t (&v)[16], R16Pass s,
                                             int64_t *lds, int c, int r) {
  s.r = r;
#pragma unroll
  for (int m = 0; m < 16; m++)
    v[m] = src[r + 16 * m];
  r16_forward_high<A>(v, s);
  r16_exchange(v, lds, c, r, true);
  r16_forward_low<A>(v, s);
  r16_forward_settle<A>(v, s);
}

// Inverse low-bit pass from COLS: bits 0..3, exchange to ROWS, bits 4..7,
// store ROWS (coalesced).
template <int A>
__device__ __forceinline__ void inverse_tile(int64_t (&v)[16], int64_t *dst,
                                             R16Pass s, int64_t *lds, int c,
                                             int r) {
  r16_inverse_low<A>(v, s);
  r16_exchange(v, lds, c, r, false);
  r16_inverse_high<A>(v, s);
#pragma unroll
  for (int m = 0; m < 16; m++)
    dst[r + 16 * m] = v[m];
}

template <int L>
__device__ __forceinline__ void tensor_tile(const TensorArgs &a,
                                            const DeviceTables &t,
                                            int64_t *lds, int ct, int tile,
                                            int l) {
  const int64_t n = int64_t(1) << t.log_n;
  const int p = a.moduli_id + l;
  const int tid = threadIdx.x;
  // ROWS for the loads: column c = tid >> 4, row r = tid & 15; the same
  // (c, r) pair is used in COLS after the exchange
  const int c = tid >> 4, r = tid & 15;
  const int64_t sp = int64_t(tile) * 16 + c;
  const int64_t off = sp << 8;
  const int64_t poly = int64_t(a.limbs) * n;

  R16Pass fw;
  fw.tw = t.root + int64_t(p) * n;
  fw.ts = t.root_shoup + int64_t(p) * n * 2;
  fw.n = n;
  fw.hi = sp;
  fw.log_n = t.log_n;
  fw.b = 0;
  fw.r = r;
  fw.q = t.q[p];
  fw.k = t.k[p];
  fw.q2 = int64_t(t.q2[p]);
  fw.nq2 = opaque_neg(fw.q2);
  fw.nq = uint64_t(opaque_neg(int64_t(fw.q)));
  R16Pass iv = fw;
  iv.tw = t.root_inverse + int64_t(p) * n;
  iv.ts = t.root_inverse_shoup + int64_t(p) * n * 2;
  const uint64_t q = fw.q, k = fw.k;
  const int64_t q2 = fw.q2;

  const int64_t *a0 = a.a + int64_t(ct) * 2 * poly + int64_t(l) * n + off;
  const int64_t *b0 = a.b + int64_t(ct) * 2 * poly + int64_t(l) * n + off;
  int64_t va0[16], vb0[16], va1[16], vb1[16];
  forward_tile<L>(a0, va0, fw, lds, c, r);
  forward_tile<L>(b0, vb0, fw, lds, c, r);
  forward_tile<kExact>(a0 + poly, va1, fw, lds, c, r);
  forward_tile<kExact>(b0 + poly, vb1, fw, lds, c, r);

  // the tensor, in the order of tensor_kernel
  int64_t z1[16];
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int64_t x = mont_mult_u(va0[m], vb1[m], q, k);
    const int64_t y = mont_mult_u(va1[m], vb0[m], q, k);
    z1[m] = guard_u(guard_u(add64(x, y), q2), int64_t(q));
    va0[m] = mont_mult_u(va0[m], vb0[m], q, k); // z0
    va1[m] = mont_mult_u(va1[m], vb1[m], q, k); // z2
  }

  int64_t *z = a.trio + int64_t(ct) * 3 * poly + int64_t(l) * n + off;
  inverse_tile<L>(va0, z, iv, lds, c, r);
  inverse_tile<L>(z1, z + poly, iv, lds, c, r);
  inverse_tile<kExact>(va1, z + 2 * poly, iv, lds, c, r);
}

__global__ __launch_bounds__(kBlock, 1) void tensor_r16_kernel(
    TensorArgs a, DeviceTables t) {
  __shared__ int64_t lds[kR16Lds];
  const int tiles = int((int64_t(1) << t.log_n) >> 12);
  int64_t w = blockIdx.x;
  if (a.swizzle) {
    const int64_t per = int64_t(gridDim.x) >> 3;
    w = int64_t(blockIdx.x & 7) * per + (blockIdx.x >> 3);
  }
  // ciphertext fastest: neighbouring blocks share the limb's twiddles
  const int ct = int(w % a.count);
  const int64_t rest = w / a.count;
  const int tile = int(rest % tiles);
  const int l = int(rest / tiles);
  if (!a.lazy)
    tensor_tile<kExact>(a, t, lds, ct, tile, l);
  else if (t.q[a.moduli_id + l] < kLazySmallBound)
    tensor_tile<kLazySmall>(a, t, lds, ct, tile, l);
  else
    tensor_tile<kLazyLarge>(a, t, lds, ct, tile, l);
}

} // namespace

bool tensor_fused_supported(const DeviceTables &t) { return t.log_n == 16; }

void launch_tensor_fused(const DeviceTables &t, const int64_t *a,
                         const int64_t *b, int64_t *trio, int count, int limbs,
                         int moduli_id, bool lazy, hipStream_t stream) {
  const int tiles = int((int64_t(1) << t.log_n) >> 12);
  const int64_t blocks = int64_t(count) * limbs * tiles;
  TensorArgs args{a, b, trio, count, limbs, moduli_id, lazy ? 1 : 0,
                  blocks % 8 == 0 ? 1 : 0};
  hipLaunchKernelGGL(tensor_r16_kernel, dim3(unsigned(blocks)), dim3(kBlock),
                     0, stream, args, t);
}

} // namespace fhe
