// Host-side launchers for the HIP kernels.  All limb arrays are device int64
// in [polynomial][limb][coefficient] order; a limb at flat index f of a batch
// whose polynomials have `limbs` limbs uses prime moduli_id + f % limbs.
#pragma once

#include "device.h"

namespace fhe {

// Which polynomials of a batched transform may use the lazy arithmetic
// (ntt_r16.h): those whose index modulo `period` is below `count`.  Lazy
// outputs are other representatives of the same residues: forward outputs
// below 2q, or below 2^48 for primes under 2^42; inverse outputs with the
// exit below 2q (canonical with kInttExitReduce).  The default is exact
// everywhere.
struct LazySpec {
  int period = 0;
  int count = 0;
  // 0: every polynomial; 1: only the exact ones; 2: only the lazy ones
  // (split per-variant launches of the radix-16 passes only)
  int part = 0;
};
constexpr LazySpec kAllLazy{1, 1};

// Whether LazySpec::part may select a subset of the polynomials: the
// radix-16 passes with per-variant launches.
bool ntt_parts_supported(const DeviceTables &t);

// In-place capable forward NTT of total_limbs limbs.  pre (device, indexed by
// absolute prime) multiplies each input value in Montgomery form first, e.g.
// R^2 for the reference's to_ntt (mont_enter then ntt); nullptr for none.
void ntt_forward(const DeviceTables &t, const int64_t *src, int64_t *dst,
                 int64_t total_limbs, int limbs, int moduli_id,
                 const int64_t *pre, hipStream_t stream,
                 LazySpec lazy = LazySpec());

// In-place capable inverse NTT with the InttPost epilogue.
void ntt_inverse(const DeviceTables &t, const int64_t *src, int64_t *dst,
                 int64_t total_limbs, int limbs, int moduli_id, int post,
                 hipStream_t stream, LazySpec lazy = LazySpec());

// Single passes of the two-pass transform (N >= 2^13): the forward pass over
// the high index bits, and the inverse pass over the high index bits with
// the InttPost epilogue.  The remaining passes run inside fused kernels.
void ntt_forward_first(const DeviceTables &t, const int64_t *src, int64_t *dst,
                       int64_t total_limbs, int limbs, int moduli_id,
                       const int64_t *pre, hipStream_t stream,
                       LazySpec lazy = LazySpec());
void ntt_inverse_second(const DeviceTables &t, const int64_t *src,
                        int64_t *dst, int64_t total_limbs, int limbs,
                        int moduli_id, int post, hipStream_t stream,
                        LazySpec lazy = LazySpec());
// The forward pass over the low index bits, after ntt_forward_first: the two
// together are ntt_forward.
void ntt_forward_second(const DeviceTables &t, const int64_t *src,
                        int64_t *dst, int64_t total_limbs, int limbs,
                        int moduli_id, hipStream_t stream,
                        LazySpec lazy = LazySpec());

// Fused key-switching core for N = 2^16: for each tile of each extended limb
// and ciphertext, the low-bit forward NTT pass of every partition's
// extended polynomial, the key inner product over all partitions
// (src/ckks/core_key_switch.cpp:318-336) and the low-bit inverse NTT pass of
// both accumulators.  ext: [count][parts][ext_limbs][N] after
// ntt_forward_first; acc: [count][2][ext_limbs][N] before ntt_inverse_second.
bool ks_fused_supported(const DeviceTables &t);
void launch_ks_inner_product_fused(const DeviceTables &t, const int64_t *ext,
                                   const int64_t *key, int key_chain,
                                   int key_offset, int key_partition0,
                                   int parts, int64_t *acc, int count,
                                   int ext_limbs, int moduli_id, bool lazy,
                                   const int64_t *own, int64_t own_stride,
                                   const int *starts, const int *sizes,
                                   bool lazy_accumulate, bool wide_accumulate,
                                   bool prefetch, int occupancy,
                                   hipStream_t stream);

// out = mont_mult(a, b) elementwise (src/montgomery/unit.cpp:5-10).
void launch_mont_mult(const DeviceTables &t, const int64_t *a,
                      const int64_t *b, int64_t *out, int64_t total_limbs,
                      int limbs, int moduli_id, hipStream_t stream);

// out = mont(a, b) with b one polynomial [limbs][N] for every polynomial of
// a [total_limbs / limbs][limbs][N] (plaintext products of a batch).
void launch_mont_mult_broadcast(const DeviceTables &t, const int64_t *a,
                                const int64_t *b, int64_t *out,
                                int64_t total_limbs, int limbs, int moduli_id,
                                hipStream_t stream);

// NTT-form tensor of `count` duo pairs -> trios (src/ckks/core.cpp:284-314).
void launch_tensor(const DeviceTables &t, const int64_t *a, const int64_t *b,
                   int64_t *trio, int count, int limbs, int moduli_id,
                   hipStream_t stream);

// Garner digits of one partition and exact basis extension to all
// limbs + alpha primes (src/ckks/core_key_switch.cpp:21-108).  x: INTT-form
// polynomial of each ciphertext at x + c * x_stride; ext: [count][ext][N].
// non_negative: the partition's first prime is at most twice each of its
// other primes (CkksParams::modup_non_negative), which keeps every Garner
// difference non-negative and allows the unsigned Montgomery product.
// lazy: the extended limbs may be any representatives (Shoup products with
// the standard-form constants at constants + ext_shoup_offset); only for
// non-negative partitions, whose digits are canonical either way.
void launch_modup(const DeviceTables &t, const int64_t *x, int64_t x_stride,
                  int64_t *ext, int64_t ext_stride, int count, int limbs,
                  int moduli_id, int start, int size, const int64_t *constants,
                  int ext_shoup_offset, bool non_negative, bool lazy,
                  hipStream_t stream);

// ModUp of every partition fused with the first forward NTT pass of every
// extended limb (N = 2^16, partitions of at most 4 limbs): ext
// [count][parts][ext][N] receives what launch_modup followed by
// ntt_forward_first would write (the same representatives with lazy = false,
// the same residues otherwise).  starts / sizes / is_signed: per partition.
constexpr int kMaxFusedParts = 32;
bool modup_ntt_fused_supported(const DeviceTables &t, int max_partition_size,
                               int parts);
void launch_modup_ntt_fused(const DeviceTables &t, const int64_t *x,
                            int64_t x_stride, int64_t *ext, int count,
                            int limbs, int moduli_id, const int *starts,
                            const int *sizes, const bool *is_signed, int parts,
                            const int64_t *constants, int constants_stride,
                            bool lazy, hipStream_t stream);

// ModUp staged through the Garner digits (N = 2^16, partitions of at most
// 4 limbs): digits [count][parts][maxz][N] (maxz = partition size rounded up
// to a power of two) then ext [count][parts][ext][N] after the first forward
// NTT pass, as launch_modup + ntt_forward_first would give (the same
// representatives without lazy; with lazy, for non-negative partitions, the
// same residues).
bool staged_modup_supported(const DeviceTables &t, int max_partition_size,
                            int parts);
void launch_staged_modup(const DeviceTables &t, const int64_t *x,
                         int64_t x_stride, int64_t *digits, int64_t *ext,
                         int count, int limbs, int moduli_id,
                         const int *starts, const int *sizes,
                         const bool *is_signed, int parts,
                         const int64_t *constants, int constants_stride,
                         int ext_shoup_offset, bool lazy, bool skip_own,
                         hipStream_t stream);

// Second forward NTT pass of both input duos, tensor and first inverse NTT
// pass of the trio in one kernel (N = 2^16).  a, b: [count][2][limbs][N]
// after ntt_forward_first; trio: [count][3][limbs][N] for
// ntt_inverse_second.  With lazy, a0 / b0 / z0 / z1 use the lazy arithmetic
// (LazySpec{2, 1} for the first passes, {3, 2} for the last).
bool tensor_fused_supported(const DeviceTables &t);
void launch_tensor_fused(const DeviceTables &t, const int64_t *a,
                         const int64_t *b, int64_t *trio, int count, int limbs,
                         int moduli_id, bool lazy, hipStream_t stream);

// Key inner product for one partition: acc[c][h] (+)= ext[c] * key_h
// (src/ckks/core_key_switch.cpp:318-336).  Each key limb is read once for
// the whole batch.  key0/key1 point at the key polynomials' first limb
// aligned with the ciphertext's moduli_id.
void launch_key_inner_product(const DeviceTables &t, const int64_t *ext,
                              int64_t ext_stride, const int64_t *key0,
                              const int64_t *key1, int64_t *acc,
                              int64_t acc_stride, int count, int ext_limbs,
                              int moduli_id, bool first, hipStream_t stream);

// ModDown of both key-switched halves, addition of the relinearized d0/d1 and
// rescale, fused per coefficient (src/ckks/core_key_switch.cpp:110-160,
// src/ckks/core.cpp:736-737, src/ckks/core_rescale.cpp:5-25).
//   acc:  [count][2][limbs+alpha][N] INTT form, reduced to [0, q)
//   trio: [count][3][limbs][N], polynomials 0/1 in INTT form
//   out:  [count][2][limbs-1][N]
void launch_moddown_relin_rescale(const DeviceTables &t, const int64_t *acc,
                                  const int64_t *trio, int64_t *out, int count,
                                  int limbs, int moduli_id, hipStream_t stream);

// ModDown + add without rescale: out[c][h] = guard(d_h + ModDown(acc_h)).
void launch_moddown_relin(const DeviceTables &t, const int64_t *acc,
                          const int64_t *trio, int64_t *out, int count,
                          int limbs, int moduli_id, hipStream_t stream);

// ModDown of both halves plus an optional addend: out[c][h] =
// guard(d_h + ModDown(acc_h)) where d_h = d + c * d_stride + h * limbs * N
// for h < d_polys and zero otherwise (the c0 term of a rotation,
// src/ckks/core_key_switch.cpp:509-548, and the copy of the second half in
// src/ckks/core.cpp:1035-1056).
// The butterfly stage with a BootstrapKey in the extended basis
// (bk_kernels.hip): E = limbs + alpha limbs per polynomial at moduli_id, NTT
// form, values kept below 2q.
constexpr int kBkMaxTerms = 16;
constexpr int kBkMaxGroups = 8;
constexpr int kBkMaxOuter = 64;
// One term of a batched ntt_rotate: key inner product of the decomposition
// dec (ciphertext c at dec + c * dec_stride, [parts][E][N]) with the key
// (half h of partition p at key + (2 p + h) * key_chain * N, limb 0 aligned
// with the ciphertext's limb 0), plus the source's polynomial 0 (c0 + c *
// src_stride) in half 0, permuted by X -> X^g; with dec == nullptr the
// identity: polynomials c0 and c1 as they are.
struct BkTerm {
  const int64_t *c0;
  const int64_t *c1;
  const int64_t *dec;
  const int64_t *key;
  int64_t src_stride;
  int64_t dec_stride;
  int key_chain;
  uint32_t g;
};
struct BkRotateArgs {
  BkTerm term[kBkMaxTerms];
  int terms;
  int count;
  int parts;
  int ext_limbs;
  int moduli_id;
  int accumulate;     // sum launches: add to out instead of overwriting
  int64_t *out;       // half 1 is E * N words after half 0
  int64_t out_c_stride;
  int64_t out_t_stride; // term launches: per term
};
// Products with the diagonals of a stage summed per group:
// sums[h][group][c] = sum over members o of the group and inner i of
// rotated_i[c][h] * diagonal[o][i] (o, i) != (0, 0); rotated_0 = base.
struct BkDiagArgs {
  const int64_t *base;       // [count][2][E][N]
  const int64_t *rotated;    // [count][inner - 1][2][E][N]
  const int64_t *const *diagonal; // device table [outer][inner]
  int64_t *sums;             // [2][groups][count][E][N]
  int inner;
  int count;
  int ext_limbs;
  int moduli_id;
  int groups;                // groups in sums
  int group0;                // this launch's groups: group0 ..
  int groups_here;           //   group0 + groups_here - 1
  int member_start[kBkMaxGroups + 1];
  int member[kBkMaxOuter];   // outer indices of the launch's groups
};
bool bk_supported(const DeviceTables &t);
void launch_bk_rotate(const DeviceTables &t, const BkRotateArgs &a, bool sum,
                      hipStream_t stream);
void launch_bk_diagonal(const DeviceTables &t, const BkDiagArgs &a,
                        hipStream_t stream);
void launch_bk_attach(const DeviceTables &t, const int64_t *in, int64_t *out,
                      int64_t polys, int limbs, int ext_limbs,
                      hipStream_t stream);

// ModDown alone (CKKSCore::mod_down after its INTT and reduction,
// src/ckks/core_key_switch.cpp:110-160) of `polys` INTT-form extended
// polynomials acc [polys][limbs + alpha][N] (values in [0, q)) into
// out [polys][limbs][N], canonical.
void launch_moddown_polys(const DeviceTables &t, const int64_t *acc,
                          int64_t *out, int64_t polys, int limbs,
                          int moduli_id, hipStream_t stream);
void launch_moddown_add(const DeviceTables &t, const int64_t *acc,
                        const int64_t *d, int64_t d_stride, int d_polys,
                        int64_t *out, int count, int limbs, int moduli_id,
                        hipStream_t stream);

// Coefficient-domain automorphism X -> X^g of total_limbs INTT-form limbs
// (g odd, below 2N), as the reference's rotate_ with its permutation and
// sign (src/ckks/core.cpp:845-869, src/ckks/encode.cpp:513-526):
// out[(g i mod 2N) mod N] = +-in[i], then make_unsigned_reduce_2q_to_q.
void launch_automorphism(const DeviceTables &t, const int64_t *in,
                         int64_t *out, int64_t total_limbs, int limbs,
                         int moduli_id, uint32_t g, hipStream_t stream);

// NTT-domain automorphism: the same map on the evaluation points of NTT-form
// limbs, a pure gather out[i] = in[perm(i)] (src/ckks/core.cpp:886-910,
// src/ckks/encode.cpp:528-557).
void launch_ntt_automorphism(const DeviceTables &t, const int64_t *in,
                             int64_t *out, int64_t total_limbs, uint32_t g,
                             hipStream_t stream);

// Per-coefficient primitives of src/montgomery/unit.cpp:5-86,
// src/montgomery/compound.cpp:6-109 and src/operation/scalar.cpp:5-11; the
// codes equal the FHE_EW_* values of the C ABI.  x2 is the second operand
// (or the accumulator of the enter-add forms), c the per-limb constant.
enum EwOp : int {
  kEwMontMult = 0,            // mont(x1, x2)
  kEwMontEnter = 1,           // mont(x1, c)
  kEwReduce = 2,              // guard(x1, q)
  kEwMakeSigned = 3,          // x1 <= q/2 (unsigned compare) ? x1 : x1 - q
  kEwNegate = 4,              // -x1
  kEwAdd = 5,                 // guard(x1 + x2, 2q)
  kEwSubtract = 6,            // x1 - x2, + 2q when negative
  kEwEnterAdd = 7,            // guard(mont(x1, c) + x2, 2q)
  kEwEnterAddReduce = 8,      // guard(guard(mont(x1, c) + x2, 2q), q)
  kEwMultAdd = 9,             // guard(mont(x1, x2) + y, 2q), y in place
  kEwSubEnterReduce = 10,     // guard(mont(guard(x1 + 2q - x2, 2q), c), q)
  kEwSubEnter = 11,           // mont(guard(x1 + 2q - x2, 2q), c)
  kEwEnterExit = 12,          // exit(mont(x1, c))
  kEwEnterExitReduce = 13,    // guard(exit(mont(x1, c)), q)
  kEwAddReduce = 14,          // guard(guard(x1 + x2, 2q), q)
  kEwSubReduce = 15,          // guard(guard(x1 + 2q - x2, 2q), q)
  kEwMakeUnsignedReduce = 16, // x1 < 0 ? x1 + q : guard(x1, q)
  kEwNegateMakeUnsigned = 17, // q - x1
  kEwAddScalar = 18,          // coefficient 0: x1 + c, others x1
  kEwRounder = 19,            // x1 > q/2 (unsigned compare) ? 1 : 0
  kEwEnterReduce = 20,        // guard(mont(x1, c), q)
  kEwAddScalarReduce = 21,    // guard(coefficient 0: x1 + c, others x1, q)
  kEwOpCount = 22
};
constexpr int kMaxEwLimbs = 128;
bool ew_needs_x2(int op);
bool ew_needs_constants(int op);
// constants: host array of `limbs` values (or nullptr)
// copy_period > 0: the operation applies to polynomial 0 of every group of
// copy_period polynomials (of `limbs` limbs each) and the others are copied
// from x1 unchanged.
void launch_elementwise(const DeviceTables &t, int op, const int64_t *x1,
                        const int64_t *x2, int64_t *y, int64_t total_limbs,
                        int limbs, int moduli_id, const int64_t *constants,
                        hipStream_t stream, int copy_period = 0);

// y = rescale(guard(mont(x, c), q)) per polynomial, dropping `skip` more
// limbs after the rescale (scalar multiply, level down).  constants: host,
// `limbs` values; limbs <= kMaxEwLimbs.
// weighted_sum of up to kWsMaxTerms texts at one level and form (not
// rescaled, coefficient domain), each with its multiplier constants per
// limb: out [polys][limbs - 1][N] equals the engine's sequence of
// multiply(basis, weight) and add calls.
constexpr int kWsMaxTerms = 8;
constexpr int kWsMaxLimbs = 32;
void launch_weighted_sum(const DeviceTables &t,
                         const int64_t *const *inputs,
                         const int64_t *const *constants, int terms,
                         int64_t *out, int polys, int limbs, int moduli_id,
                         hipStream_t stream);
void launch_scalar_rescale(const DeviceTables &t, const int64_t *in,
                           int64_t *out, int polys, int limbs, int moduli_id,
                           const int64_t *constants, int skip,
                           hipStream_t stream);

// Encryption arithmetic: samples [count][N] lifted to [count][limbs][N]
// with make_unsigned_reduce_2q_to_q, and the error / plaintext fold of
// CKKSCore::encrypt (src/ckks/core.cpp:35-82) into out [count][2][limbs][N].
void launch_tile_unsigned(const DeviceTables &t, const int64_t *samples,
                          int64_t *out, int count, int limbs, int moduli_id,
                          hipStream_t stream);
// the monomial value * X^pos tiled over `limbs` limbs from moduli_id
// (|value| < every q)
void launch_tile_monomial(const DeviceTables &t, int64_t value, int64_t pos,
                          int64_t *out, int limbs, int moduli_id,
                          hipStream_t stream);
void launch_encrypt_finish(const DeviceTables &t, const int64_t *pt,
                           const int64_t *e0, const int64_t *e1, int64_t *out,
                           int count, int limbs, int moduli_id,
                           hipStream_t stream);

// Encoding tables on the device (host-built, see CkksParams::encode_*).
struct EncodeDevice {
  const uint32_t *pre;      // slot -> codec index
  const double *fft_cos, *fft_sin;
  const double *twist_cos, *twist_sin;
  int log_codec;
};
// Encode `count` messages x ([count][slots] doubles, or [count][slots][2]
// for complex input) into out [count][out_limbs][N]: limbs [0, first) are
// zero (the top level of the one-more-level chain), limbs from `first` use
// primes moduli_id + (l - first) and the constants r_mult (host).  work:
// encode_work_bytes(count) device bytes.
size_t encode_work_bytes(const EncodeDevice &e, int count);
void launch_encode(const DeviceTables &t, const EncodeDevice &e,
                   const double *x, int complex_input, int count,
                   double scale_before_tile, int out_limbs, int first,
                   int moduli_id, const int64_t *r_mult, int64_t *out,
                   void *work, hipStream_t stream);

// Decoding tables on the device (host-built, see CkksParams::decode_*).
struct DecodeDevice {
  const uint32_t *post;     // codec index -> slot
  const double *skew_cos, *skew_sin;
  const double *ifft_cos, *ifft_sin;
  int log_codec;
};
// CRT constants of one moduli_id (CkksParams::DecodeCrt on the device).
struct DecodeCrtDevice {
  const uint64_t *reverser; // [limbs][words]
  const uint64_t *q_product, *q_half, *upper;
  uint64_t k;
  int words;
};
constexpr int kMaxDecodeWords = 32;
bool decode_crt_supported(int words);
size_t decode_work_bytes(const DecodeDevice &d, int count);
// Decode `count` INTT-form plaintexts in [count][limbs][N] at moduli_id into
// out: [count][slots] real parts, or [count][slots][2] when complex_output.
void launch_decode(const DeviceTables &t, const DecodeDevice &d,
                   const DecodeCrtDevice &crt, const int64_t *in, int count,
                   int limbs, int moduli_id, double scale_inverse,
                   int complex_output, double *out, void *work,
                   hipStream_t stream);

// Modular products per second of one form (0: Shoup, 1: Montgomery) on the
// device, from a kernel of independent product chains over `blocks`
// workgroups (out: blocks * kBlock words); synchronizes the stream.  A
// negative value reports a runtime failure.
double measure_modmul_rate(int form, int blocks, int iters, uint64_t q,
                           uint64_t k, uint64_t w, uint64_t wq, uint64_t *out,
                           hipStream_t stream);

// Rescale of `polys` INTT-form polynomials (src/ckks/core_rescale.cpp:5-25).
void launch_rescale(const DeviceTables &t, const int64_t *in, int64_t *out,
                    int polys, int limbs, int moduli_id, hipStream_t stream);

} // namespace fhe
