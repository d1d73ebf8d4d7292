/*
 * fhe_amd.h -- C ABI of the MI355X CKKS evaluation backend.
 *
 * A drop-in device backend for the evaluation hot path of the reference CKKS
 * engine (Desilo FHE).  The reference has no C ABI or plugin registry: its
 * backends are the closed enum EngineProcessor {CPU, CPU_PARALLEL}
 * (src/engine/processor.h:5-8) behind AbstractEngine
 * (src/engine/abstract_engine.h:12-299).  These entry points are what a host
 * engine binds in place of that processor seam; every one names the
 * reference interface whose work it replaces.
 *
 * Conventions
 *  - Ciphertext / key / polynomial data are int64 limb arrays in the
 *    reference's order [polynomial][limb][coefficient] (one limb = N int64).
 *    Limb l of an object at moduli_id m is a residue modulo prime m + l of the
 *    chain (src/ckks/precomputed.cpp:227-230); special primes follow the Q
 *    chain.  "Device" arrays live in GPU memory (fhe_device_alloc).
 *  - INTT-form data is in the coefficient domain; NTT-form data is
 *    Montgomery-entered (x * 2^62 mod q) in the evaluation domain, as in the
 *    reference (src/ckks/core_ntt.cpp:91-150).
 *  - Every call returns FHE_OK or an error code; fhe_last_error() gives the
 *    thread's last message.  Validation failures return before any device
 *    work is queued.
 *  - Calls on one context are ordered on that context's HIP stream and are
 *    not thread-safe; use one context per thread.  Work is asynchronous with
 *    respect to the host until fhe_synchronize or a copy to the host.
 */
#ifndef FHE_AMD_H
#define FHE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FHE_OK 0
#define FHE_ERR_INVALID 1     /* bad argument / incompatible shapes */
#define FHE_ERR_DEVICE 2      /* HIP runtime failure */
#define FHE_ERR_OUT_OF_MEMORY 3
#define FHE_ERR_UNSUPPORTED 4

typedef struct fhe_context fhe_context;

/* Mirrors the fields of CKKSConfig (src/ckks/config.h:11-34) and the scheme
 * preset tuple (src/engine/scheme_preset.cpp:8-16) that fix the prime chain. */
typedef struct fhe_params {
  int32_t log_coeff_count;     /* log2 N: 3, 4, 13..17 */
  int32_t special_prime_count; /* alpha, 1..16 */
  int32_t scale_bits;          /* 40 (CKKS_40_60_DEFAULT) or 54 (bootstrap) */
  int32_t small_prime_bits;    /* 40 or 54 */
  int32_t one_more_level;      /* 1 for every reference preset */
  int32_t max_level;           /* -1: full level */
  int32_t prime_count;         /* >0: use primes[] instead of the searches */
  const uint64_t *primes;      /* Q chain then specials, all < 2^61 */
} fhe_params;

/* Host only (no device work): the prime chain a context with these
 * parameters uses, i.e. CKKSParameters::get_moduli(0)
 * (src/ckks/parameters.cpp:9-51 with src/arithmetic/prime.cpp:974-1015). */
int fhe_params_resolve_primes(const fhe_params *params, uint64_t *primes,
                              int capacity, int *prime_count);

/* Context = CKKSParameters + CKKSPrecomputed on the device
 * (src/ckks/parameters.cpp:9-105, src/ckks/precomputed.cpp:33-217). */
int fhe_context_create(const fhe_params *params, fhe_context **out);
void fhe_context_destroy(fhe_context *ctx);
const char *fhe_last_error(void);

/* info[0..] = N, prime_count, chain_count, special_prime_count, full_level,
 * max_level, max_key_switching_level, slot_count, scale_bits */
#define FHE_INFO_COUNT 9
int fhe_context_info(const fhe_context *ctx, int64_t *info, int count);
/* the full prime chain q[0..prime_count) (CKKSPrecomputed::get_q(0)) */
int fhe_context_primes(const fhe_context *ctx, uint64_t *out, int capacity);
/* per-prime constants {k, R^2 mod q, N^-1 R mod q} and the N-1 Montgomery
 * twiddles of each direction (src/ckks/derived_parameters.cpp:19-50) */
int fhe_context_prime_constants(const fhe_context *ctx, int prime_index,
                                int64_t *constants3, int64_t *root,
                                int64_t *root_inverse);
/* the context's hipStream_t, for event timing by the caller */
void *fhe_context_stream(fhe_context *ctx);
int fhe_synchronize(fhe_context *ctx);

/* Device memory. */
int fhe_device_alloc(fhe_context *ctx, size_t bytes, void **out);
int fhe_device_free(fhe_context *ctx, void *ptr);
int fhe_copy_to_device(fhe_context *ctx, void *dst, const void *src,
                       size_t bytes);
int fhe_copy_to_host(fhe_context *ctx, void *dst, const void *src,
                     size_t bytes);

/* Forward NTT of polys x limbs device limbs in place.
 * enter = 1: CKKSCore::to_ntt_ (mont_enter(R^2) then ntt,
 *            src/ckks/core_ntt.cpp:91-100);
 * enter = 0: the bare transform (src/operation/ntt.cpp:46-134). */
int fhe_ntt(fhe_context *ctx, int64_t *data, int polys, int limbs,
            int moduli_id, int enter);

/* Inverse NTT in place.  post = 0 bare (src/operation/ntt.cpp:177-272),
 * 1 CKKSCore::to_intt_ (intt + mont_enter_exit(N^-1 R),
 * src/ckks/core_ntt.cpp:131-150), 2 as 1 then reduce_2q_to_q. */
int fhe_intt(fhe_context *ctx, int64_t *data, int polys, int limbs,
             int moduli_id, int post);

/* out = a * b * 2^-62 elementwise (SingleProcessor::mont_mult,
 * src/processor/single.cpp:60-74). */
int fhe_mont_mult(fhe_context *ctx, const int64_t *a, const int64_t *b,
                  int64_t *out, int polys, int limbs, int moduli_id);

/* Rescale of polys INTT-form polynomials at moduli_id into polys
 * polynomials at moduli_id + 1 (CKKSCore::rescale_once_,
 * src/ckks/core_rescale.cpp:5-25). */
int fhe_rescale(fhe_context *ctx, const int64_t *in, int64_t *out, int polys,
                int moduli_id);

/* Relinearization of `count` NTT-form trios [count][3][limbs][N] at
 * moduli_id with a relinearization key [key_polys][key_limbs][N] at
 * key_moduli_id (device) into INTT-form duos [count][2][limbs][N]
 * (CKKSCore::relinearize, src/ckks/core.cpp:713-741). */
int fhe_relinearize(fhe_context *ctx, int count, int moduli_id,
                    const int64_t *trios, const int64_t *key,
                    int key_moduli_id, int64_t *out);

/* The metric's unit of work for `count` ciphertext pairs in one batch:
 * CKKSEngine::multiply(ct, ct, rlk) followed by CKKSEngine::rescale(ct)
 * (src/ckks/engine_multiply.cpp:149-162, src/ckks/engine_rescale.cpp:50-60)
 * on rescaled INTT-form duos a, b [count][2][limbs][N] at moduli_id.  The key
 * is streamed once per batch.  out: [count][2][limbs-1][N] at moduli_id + 1,
 * INTT form. */
int fhe_hmult_relin_rescale(fhe_context *ctx, int count, int moduli_id,
                            const int64_t *a, const int64_t *b,
                            const int64_t *key, int key_moduli_id,
                            int64_t *out);

/* No counterpart in the reference: the same computation as
 * fhe_hmult_relin_rescale, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation).  It calls no hipMalloc / hipFree: its workspace is the
 * library's fixed device scratch array of 512 MiB per GPU (held until the
 * process ends), and a batch whose workspace exceeds it is refused with
 * FHE_ERR_INVALID before any launch (a Gold ciphertext pair at level 24 needs
 * ~282 MB).  The key must be at the maximum key-switching level (whose ModUp
 * constants the context prepares at creation).  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (two contexts' streams,
 * two host threads) race. */
int fhe_hmult_relin_rescale_noalloc(fhe_context *ctx, int count,
                                    int moduli_id, const int64_t *a,
                                    const int64_t *b, const int64_t *key,
                                    int key_moduli_id, int64_t *out);

/* Data flags of a text, the reference's DataFlag values (src/data/flag.h:7-14).
 * A text at level L sits at moduli_id full_level - L, or one higher when
 * rescaled. */
enum { FHE_FLAG_NTT_FORM = 2, FHE_FLAG_RESCALED = 4 };

/* CKKSEngine::multiply(ct, ct) and, with a key, multiply(ct, ct, rlk)
 * (src/ckks/engine_multiply.cpp:39-57, 134-147) of `count` duos in any form:
 * a [count][2][limbs][N] at (a_level, a_flags), b likewise.  Both go through
 * prepare_for_multiply (src/ckks/engine_multiply.cpp:13-27) at the lower
 * level L: INTT when the level changes, level down (plain or after rescale),
 * rescale, NTT.  key = NULL: out is the NTT-form trio [count][3][limbs][N];
 * with a relinearization key at key_moduli_id: the INTT-form duo
 * [count][2][limbs][N].  Either is not rescaled at level L - 1.  a == b with
 * equal forms gives square(ct) (src/ckks/engine_multiply.cpp:86-99).
 * Levels must be positive. */
int fhe_multiply(fhe_context *ctx, int count, const int64_t *a, int a_level,
                 int a_flags, const int64_t *b, int b_level, int b_flags,
                 const int64_t *key, int key_moduli_id, int64_t *out);

/* CKKSEngine::multiply(ct, pt) (src/ckks/engine_multiply.cpp:210-227) of
 * `count` duos or trios ct [count][ct_polys][limbs][N] with one plaintext
 * each, pt [count][limbs][N], both in any form: prepared as in
 * fhe_multiply, then every polynomial times the plaintext
 * (CKKSCore::multiply, src/ckks/core.cpp:257-282).  out: NTT form
 * [count][ct_polys][limbs][N], not rescaled at level min - 1. */
int fhe_multiply_plain(fhe_context *ctx, int count, const int64_t *ct,
                       int ct_polys, int ct_level, int ct_flags,
                       const int64_t *pt, int pt_level, int pt_flags,
                       int64_t *out);

/* CKKSEngine::add(ct, ct) / subtract(ct, ct) (src/ckks/engine_add.cpp:82-110,
 * src/ckks/engine_subtract.cpp:5-35) of `count` pairs of duos or trios in
 * any form, a [count][a_polys][limbs][N] at (a_level, a_flags), b likewise;
 * subtract = 1 gives a - b.  Both go through match_form
 * (src/ckks/engine_add.cpp:37-74): the first text's NTT flag decides the
 * domain, the higher text leaves the NTT domain and goes down to the lower
 * level, a rescale where only one of them is rescaled.  out
 * [count][max(a_polys, b_polys)][limbs][N] at the lower level, in a's
 * domain, rescaled if either matched input is. */
int fhe_add_texts(fhe_context *ctx, int count, const int64_t *a,
                  int a_polys, int a_level, int a_flags, const int64_t *b,
                  int b_polys, int b_level, int b_flags, int subtract,
                  int64_t *out);

/* CKKSEngine::round (src/ckks/engine_bootstrap.cpp:672-691), the bootstrap's
 * EvalMod: x * 28 - modulo(x), modulo = three double-angle steps
 * (:640-659) after the Chebyshev / Paterson-Stockmeyer sine (:459-638),
 * composed of the engine operations above with the reference's operation
 * order.  in: `count` INTT-form, not rescaled duos [count][2][limbs][N] at
 * `level`; the relinearization key at key_moduli_id.  out: room for the
 * input's size, the results packed as [count][2][limbs'][N]; the
 * result's level and data flags come back in *out_level, *out_flags (its
 * limbs are those of that form). */
int fhe_round(fhe_context *ctx, int count, const int64_t *in, int level,
              const int64_t *key, int key_moduli_id, int64_t *out,
              int *out_level, int *out_flags);

/* The keys of a bootstrap with a RotationKey, all on the device: the
 * relinearization and conjugation keys, the RotationKey's fixed rotation
 * keys in its own order (src/data/data.h:163-174) with their deltas, and the
 * SmallBootstrapKey's two key-switching keys (src/data/bootstrap.cpp). */
typedef struct fhe_bootstrap_keys {
  const int64_t *relin;
  int relin_moduli_id;
  const int64_t *conjugation;
  int conjugation_moduli_id;
  int rotation_count;
  const int64_t *const *rotation;
  const int *rotation_moduli_ids;
  const int *rotation_deltas;
  const int64_t *to_sparse;
  int to_sparse_moduli_id;
  const int64_t *to_dense;
  int to_dense_moduli_id;
} fhe_bootstrap_keys;

/* Host only: the merged butterfly diagonals of one stage
 * (get_coeff_to_slot_twiddles / get_slot_to_coeff_twiddles,
 * src/operation/bootstrap.cpp:125-263) as {re, im} doubles, *count vectors
 * of 2^log_slots entries each. */
int fhe_bootstrap_twiddles(int log_slots, int stage_count, int to_slot,
                           int stage, double *out, int64_t capacity,
                           int64_t *count);

/* CKKSEngine::coeff_to_slot with a RotationKey
 * (src/ckks/engine_bootstrap.cpp:330-366): stage_count butterfly stages
 * (baby-step rotations, products with the merged diagonals of
 * src/operation/bootstrap.cpp:125-218, giant-step rotations), then the real
 * and imaginary parts by conjugation.  in: `count` duos [count][2][limbs][N]
 * at (level, flags); real and imag receive `count` duos each, packed (room
 * for the input's size), both of the form returned in *out_level,
 * *out_flags.  The ciphertexts share every key read and every diagonal. */
int fhe_coeff_to_slot(fhe_context *ctx, int count, const int64_t *in,
                      int level, int flags, const fhe_bootstrap_keys *keys,
                      int stage_count, int64_t *real, int64_t *imag,
                      int *out_level, int *out_flags);

/* CKKSEngine::slot_to_coeff with a RotationKey
 * (src/ckks/engine_bootstrap.cpp:405-434): real + i * imag, then the
 * inverse butterfly stages (src/operation/bootstrap.cpp:220-263).  real and
 * imag share the form (level, flags). */
int fhe_slot_to_coeff(fhe_context *ctx, int count, const int64_t *real,
                      const int64_t *imag, int level, int flags,
                      const fhe_bootstrap_keys *keys, int stage_count,
                      int64_t *out, int *out_level, int *out_flags);

/* CKKSEngine::bootstrap with a RotationKey and a SmallBootstrapKey
 * (src/ckks/engine_bootstrap.cpp:712-754): level down to 0, mod_raise,
 * (partial sums below full slots), coeff_to_slot, round of both parts,
 * slot_to_coeff, subtraction from the raised text and the normalizing
 * integer product.  in: `count` INTT-form duos [count][2][limbs][N] at
 * (level, flags), bootstrapped together (each as the reference bootstraps
 * it alone); out: room for [count][2][raised limbs][N], the results packed;
 * their form comes back. */
int fhe_bootstrap(fhe_context *ctx, int count, const int64_t *in, int level,
                  int flags,
                  const fhe_bootstrap_keys *keys, int stage_count,
                  int64_t *out, int *out_level, int *out_flags);

/* A BootstrapKey (src/data/bootstrap.h:24-44, built by
 * CKKSEngine::create_bootstrap_key, src/ckks/engine.cpp:162-222) whose
 * parts are already on the device.  Each butterfly stage is a PlainMatrix
 * (src/data/plain_matrix.h:10-24): its level, loop counts and deltas, and the
 * encoded diagonals (make_bootstrap_stage, src/ckks/engine_bootstrap.cpp:
 * 8-69: encoded at the level with the special primes, rescaled, NTT form;
 * each [limbs + alpha][N] at the level's rescaled moduli_id, limbs that of
 * the level after rescaling), given as a host array of outer_count x
 * inner_count device pointers [outer][inner] whose entry (0, 0) is unused.
 * The fixed rotation keys are the key's rotation contexts by (normalized)
 * delta, each [key polys][key limbs][N] at its moduli_id; the NTT-domain
 * permutations are computed on the device from the deltas. */
typedef struct fhe_plain_matrix {
  int level;
  int inner_count;
  int outer_count;
  const int *inner_deltas;          /* inner_count values, the first 0 */
  const int *outer_deltas;          /* outer_count values */
  const int64_t *const *diagonals;  /* outer_count * inner_count */
} fhe_plain_matrix;

/* Host only: the shapes of a BootstrapKey as CKKSEngine::
 * create_bootstrap_key builds them for a stage count and inner-loop shift
 * (size "medium" 0, "large" 1; src/ckks/engine.cpp:162-222,
 * src/ckks/engine_bootstrap.cpp:8-123): one stage's level, loop counts and
 * deltas (to_slot 1 coeff-to-slot, 0 slot-to-coeff; the delta arrays may be
 * NULL), and the rotation contexts in creation order, each delta with the
 * moduli_id of the level its key is made at.  Stage counts the reference
 * refuses are refused. */
int fhe_bootstrap_key_stage_layout(const fhe_params *params, int stage_count,
                                   int inner_loop_shift, int to_slot,
                                   int stage, int *level, int *inner_count,
                                   int *outer_count, int *inner_deltas,
                                   int *outer_deltas, int capacity);
int fhe_bootstrap_key_rotation_layout(const fhe_params *params,
                                      int stage_count, int inner_loop_shift,
                                      int *deltas, int *moduli_ids,
                                      int capacity, int *count);

/* make_bootstrap_stage (src/ckks/engine_bootstrap.cpp:8-69) on the device:
 * the encoded diagonals of one stage of a BootstrapKey with the given stage
 * count and inner-loop shift - each rotated twiddle vector encoded at the
 * stage's level over the chain limbs and the special primes, rescaled and
 * brought to NTT form - into out [count][limbs + alpha][N] in (outer,
 * inner) order without (0, 0), at the level's rescaled moduli_id.  With out
 * NULL only *count is set.  The limbs equal the reference key's. */
int fhe_bootstrap_key_encode_stage(fhe_context *ctx, int stage_count,
                                   int inner_loop_shift, int to_slot,
                                   int stage, int64_t *out,
                                   int64_t capacity_words, int *count);

typedef struct fhe_bootstrap_key_parts {
  int stage_count;
  const fhe_plain_matrix *coeff_to_slot; /* stage_count, or NULL */
  const fhe_plain_matrix *slot_to_coeff; /* stage_count, or NULL */
  int rotation_count;
  const int *rotation_deltas;
  const int64_t *const *rotation_keys;
  const int *rotation_moduli_ids;
  const int64_t *to_sparse;              /* the SmallBootstrapKey */
  int to_sparse_moduli_id;
  const int64_t *to_dense;
  int to_dense_moduli_id;
} fhe_bootstrap_key_parts;

/* A device BootstrapKey of a context: it keeps the caller's device pointers
 * (the caller keeps those buffers alive) and small tables of its own.
 * Creation checks the stages' shapes, that every delta has a key, and
 * uploads the ModUp constants of every stage level.  Destroy it before its
 * context. */
typedef struct fhe_bootstrap_key fhe_bootstrap_key;
int fhe_bootstrap_key_create(fhe_context *ctx,
                             const fhe_bootstrap_key_parts *parts,
                             fhe_bootstrap_key **out);
void fhe_bootstrap_key_destroy(fhe_bootstrap_key *bk);

/* CKKSEngine::coeff_to_slot with a BootstrapKey
 * (src/ckks/engine_bootstrap.cpp:368-398): the key's coeff-to-slot stages
 * (apply_butterfly with a PlainMatrix, :218-289: one decomposition shared by
 * the baby steps, rotations kept in the extended basis, one ModDown per
 * stage), then the real and imaginary parts by conjugation.  Arguments as
 * fhe_coeff_to_slot. */
int fhe_coeff_to_slot_bk(fhe_context *ctx, int count, const int64_t *in,
                         int level, int flags, const int64_t *conj,
                         int conj_moduli_id, const fhe_bootstrap_key *bk,
                         int64_t *real, int64_t *imag, int *out_level,
                         int *out_flags);

/* CKKSEngine::slot_to_coeff with a BootstrapKey
 * (src/ckks/engine_bootstrap.cpp:436-454).  Arguments as
 * fhe_slot_to_coeff. */
int fhe_slot_to_coeff_bk(fhe_context *ctx, int count, const int64_t *real,
                         const int64_t *imag, int level, int flags,
                         const int64_t *conj, int conj_moduli_id,
                         const fhe_bootstrap_key *bk, int64_t *out,
                         int *out_level, int *out_flags);

/* CKKSEngine::bootstrap with a BootstrapKey
 * (src/ckks/engine_bootstrap.cpp:794-834): as fhe_bootstrap with the key's
 * stages, rotations and SmallBootstrapKey. */
int fhe_bootstrap_bk(fhe_context *ctx, int count, const int64_t *in, int level,
                     int flags, const int64_t *relin, int relin_moduli_id,
                     const int64_t *conj, int conj_moduli_id,
                     const fhe_bootstrap_key *bk, int64_t *out, int *out_level,
                     int *out_flags);

/* Per-coefficient primitives over polys x limbs device limbs at moduli_id:
 * y = op(x1, x2, c[limb]) for the reference's limb operations
 * (src/montgomery/unit.cpp:5-86, src/montgomery/compound.cpp:6-109,
 * src/operation/scalar.cpp:5-11).  x2 is required by the two-operand codes,
 * `constants` (host, `limbs` values, limb l of every polynomial) by the codes
 * that use c.  y may alias x1 or x2; FHE_EW_MULT_ADD reads y as well. */
enum {
  FHE_EW_MONT_MULT = 0,              /* mont(x1, x2)                   */
  FHE_EW_MONT_ENTER = 1,             /* mont(x1, c)                    */
  FHE_EW_REDUCE_2Q_TO_Q = 2,         /* guard(x1, q)                   */
  FHE_EW_MAKE_SIGNED = 3,            /* x1 <= q/2 (unsigned) ? x1 : x1 - q */
  FHE_EW_NEGATE = 4,                 /* -x1                            */
  FHE_EW_ADD = 5,                    /* guard(x1 + x2, 2q)             */
  FHE_EW_SUBTRACT = 6,               /* x1 - x2 (+2q if negative)      */
  FHE_EW_ENTER_ADD = 7,              /* guard(mont(x1, c) + x2, 2q)    */
  FHE_EW_ENTER_ADD_REDUCE = 8,       /* ... then guard(., q)           */
  FHE_EW_MULT_ADD = 9,               /* guard(mont(x1, x2) + y, 2q)    */
  FHE_EW_SUB_ENTER_REDUCE = 10,      /* guard(mont(guard(x1+2q-x2, 2q), c), q) */
  FHE_EW_SUB_ENTER = 11,             /* mont(guard(x1 + 2q - x2, 2q), c) */
  FHE_EW_ENTER_EXIT = 12,            /* exit(mont(x1, c))              */
  FHE_EW_ENTER_EXIT_REDUCE = 13,     /* guard(exit(mont(x1, c)), q)    */
  FHE_EW_ADD_REDUCE = 14,            /* guard(guard(x1 + x2, 2q), q)   */
  FHE_EW_SUB_REDUCE = 15,            /* guard(guard(x1 + 2q - x2, 2q), q) */
  FHE_EW_MAKE_UNSIGNED_REDUCE = 16,  /* x1 < 0 ? x1 + q : guard(x1, q) */
  FHE_EW_NEGATE_MAKE_UNSIGNED = 17,  /* q - x1                         */
  FHE_EW_ADD_SCALAR = 18,            /* coefficient 0 only: x1 + c     */
  FHE_EW_ROUNDER = 19,               /* x1 > q/2 (unsigned) ? 1 : 0    */
  FHE_EW_ENTER_REDUCE = 20,          /* guard(mont(x1, c), q)          */
  FHE_EW_ADD_SCALAR_REDUCE = 21      /* guard(x1 (+ c at coefficient 0), q) */
};
int fhe_elementwise(fhe_context *ctx, int op, const int64_t *x1,
                    const int64_t *x2, int64_t *y, int polys, int limbs,
                    int moduli_id, const int64_t *constants);

/* Host only (no device work): the per-limb scalars of CKKSPrecomputed's
 * multipliers (src/ckks/precomputed.cpp:418-533), from the reference's
 * scale deviations (src/ckks/parameters.cpp:88-104) and BigInt reductions
 * (src/arithmetic/arithmetic.cpp:60-95, src/arithmetic/big_int.cpp:29-42).
 * The scalars feed fhe_elementwise: limb l of the object at the multiplier's
 * moduli_id.  `level` is a level, or a moduli_id for MULT_INT. */
enum {
  FHE_MULTIPLIER_ADD = 0,        /* make_add_multiplier(value, level)      */
  FHE_MULTIPLIER_MULT = 1,       /* make_mult_multiplier(value, level)     */
  FHE_MULTIPLIER_MULT_INT = 2,   /* make_mult_multiplier(int_value, mid)   */
  FHE_MULTIPLIER_LEVEL_DOWN = 3, /* make_level_down_multiplier(level, level2) */
  FHE_MULTIPLIER_LEVEL_DOWN_AFTER_RESCALE = 4
};
int fhe_multiplier(const fhe_context *ctx, int kind, double value,
                   int64_t int_value, int level, int level2, int64_t *out,
                   int capacity, int *count);
/* The same from parameters alone, without a context or device. */
int fhe_params_multiplier(const fhe_params *params, int kind, double value,
                          int64_t int_value, int level, int level2,
                          int64_t *out, int capacity, int *count);

/* Scalar operations on INTT-form, not-rescaled polynomials at a level
 * (moduli_id = full_level - level), `polys` polynomials of that level's
 * limbs, with the reference's multiplier constants:
 *  fhe_multiply_double: CKKSCore::multiply(ct, double)
 *    (src/ckks/core.cpp:196-226): scalar enter, reduce, rescale; out has one
 *    limb less (moduli_id + 1).
 *  fhe_multiply_int: CKKSCore::multiply(ct, int64_t) (src/ckks/core.cpp:
 *    228-251) at any moduli_id; out has the same limbs.
 *  fhe_add_double: CKKSCore::add(ct, double) (src/ckks/core.cpp:340-373) on
 *    `count` ciphertexts of `polys` polynomials: the scalar joins the
 *    constant coefficient of polynomial 0, the others are copied.
 *  fhe_level_down: CKKSCore::level_down (src/ckks/core.cpp:792-838) or, with
 *    rescaled = 1, level_down_after_rescale (:743-779), whose input sits at
 *    moduli_id + 1; out has the limbs of dst_level. */
int fhe_multiply_double(fhe_context *ctx, const int64_t *in, int64_t *out,
                        int polys, int level, double value);
int fhe_multiply_int(fhe_context *ctx, const int64_t *in, int64_t *out,
                     int polys, int moduli_id, int64_t value);
int fhe_add_double(fhe_context *ctx, const int64_t *in, int64_t *out,
                   int count, int polys, int level, double value);
int fhe_level_down(fhe_context *ctx, const int64_t *in, int64_t *out,
                   int polys, int level, int rescaled, int dst_level);

/* Public-key encryption of `count` INTT-form plaintexts pt
 * [count][limbs][N] at moduli_id with the public key pk [2][pk_limbs][N]
 * (NTT form, from pk_moduli_id, device) and the host-drawn samples u
 * (binary), e0, e1 [count][N] (small signed integers, device) into INTT-form
 * duos out [count][2][limbs][N] (CKKSCore::encrypt(pt, pk),
 * src/ckks/core.cpp:35-82; the samples are what get_binary and
 * get_single_error draw, src/ckks/core_random.cpp:7-74). */
int fhe_encrypt_pk(fhe_context *ctx, int count, int moduli_id,
                   const int64_t *pt, const int64_t *pk, int pk_moduli_id,
                   const int64_t *u, const int64_t *e0, const int64_t *e1,
                   int64_t *out);

/* Decryption of `count` INTT-form duos to plaintext limbs [count][limbs][N]
 * with the secret key sk [sk_limbs][N] (NTT form, from sk_moduli_id)
 * (CKKSCore::decrypt_duo, src/ckks/core.cpp:171-194). */
int fhe_decrypt(fhe_context *ctx, int count, int moduli_id, const int64_t *ct,
                const int64_t *sk, int sk_moduli_id, int64_t *out);

/* Automorphism X -> X^g, g = 5^-delta mod 2N (g = -1 for delta = -1, the
 * conjugation), of polys x limbs limbs at moduli_id into out (no aliasing).
 * ntt_form = 0: coefficient-domain permutation with signs and
 * make_unsigned_reduce_2q_to_q (CKKSCore::rotate_intt_form,
 * src/ckks/core.cpp:871-884, tables src/ckks/encode.cpp:513-526);
 * ntt_form = 1: the slot permutation of NTT-form data
 * (CKKSCore::rotate_ntt_form, src/ckks/core.cpp:886-914, table
 * src/ckks/encode.cpp:528-557). */
int fhe_automorphism(fhe_context *ctx, const int64_t *in, int64_t *out,
                     int polys, int limbs, int moduli_id, int delta,
                     int ntt_form);

/* Rotation of `count` INTT-form duos [count][2][limbs][N] at moduli_id by
 * the fixed rotation key made for `delta` ([key_polys][key_limbs][N] at
 * key_moduli_id, device) into INTT-form duos at moduli_id
 * (CKKSEngine::rotate(UnitCiphertext*, FixedRotationKey*),
 * src/ckks/engine_rotate.cpp:55-76; CKKSCore::rotate,
 * src/ckks/core.cpp:938-962).  out may alias in. */
int fhe_rotate(fhe_context *ctx, int count, int moduli_id, const int64_t *in,
               const int64_t *key, int key_moduli_id, int delta,
               int64_t *out);

/* Hoisted rotations: the same `count` INTT-form duos rotated by key_count
 * fixed rotation keys (keys[k] at key_moduli_ids[k], made for deltas[k])
 * into outs[k] [count][2][limbs][N].  The key-independent decomposition of
 * c1 (ModUp and the first NTT pass) is done once
 * (CKKSEngine::rotate_batch, src/ckks/engine_rotate.cpp:262-300, with
 * CKKSCore::decompose and rotate(ct, decomposed, key),
 * src/ckks/core_key_switch.cpp:434-460, src/ckks/core.cpp:938-951).  Each
 * outs[k] equals fhe_rotate with keys[k]; no output may alias the input. */
int fhe_rotate_batch(fhe_context *ctx, int count, int moduli_id,
                     const int64_t *in, int key_count,
                     const int64_t *const *keys, const int *key_moduli_ids,
                     const int *deltas, int64_t *const *outs);

/* A RotationKey on the device (src/data/data.h:163-174): its fixed rotation
 * keys in the key's order, each at its moduli_id, made for its delta. */
typedef struct fhe_rotation_key {
  int count;
  const int64_t *const *keys;
  const int *moduli_ids;
  const int *deltas;
} fhe_rotation_key;

/* CKKSEngine::evaluate_polynomial(ct, coefficients, rlk) with double
 * coefficients (src/ckks/engine_advanced.cpp:365-387): the power basis
 * x, x^2, ... (squares, then products of a power of two and the rest,
 * :22-57) and the weighted sum of the basis with the coefficients
 * (:228-269), for `count` INTT-form, not rescaled duos in [count][2][limbs][N]
 * at `level`.  out: room for the input's size, the results packed; their
 * level and flags come back.  The reference's requirements and messages. */
int fhe_evaluate_polynomial(fhe_context *ctx, int count, const int64_t *in,
                            int level, const double *coefficients,
                            int coefficient_count, const int64_t *rlk,
                            int rlk_moduli_id, int64_t *out, int *out_level,
                            int *out_flags);

/* CKKSEngine::sum_1d_by_row(ct, RotationKey) (src/ckks/engine_advanced.cpp:
 * 549-577) of `count` INTT-form, not rescaled duos at `level`: rotate by 1,
 * 2, 4, ... below the slot count and add, so every slot holds the sum of
 * all slots.  out as fhe_evaluate_polynomial. */
int fhe_sum_1d_by_row(fhe_context *ctx, int count, const int64_t *in,
                      int level, int flags, const fhe_rotation_key *rk,
                      int64_t *out, int *out_level, int *out_flags);

/* Serving (no counterpart in the reference, which runs one call at a time).
 * fhe_serve runs a request loop on the calling thread until every job is
 * done: job j arrives jobs[j].arrival_ms after the call starts and waits in
 * a queue per (engine, kind, level); whenever an engine has fewer than
 * max_in_flight batches on the device, its queue with the earliest waiting
 * job gives up to max_batch jobs to one batch (continuous batching: jobs
 * that arrive while others run join the next batch).  A batch is admitted
 * when the free HBM minus its estimated working set stays above
 * reserve_bytes, else it waits for a running batch to finish.  Every engine
 * (a context with its keys, all resident on the device) has its own
 * stream, so batches of different engines overlap on the device.  Job
 * inputs and outputs are device buffers; each job's output limbs equal the
 * composition run alone (fhe_multiply + fhe_sum_1d_by_row, or
 * fhe_evaluate_polynomial).  done_ms is when the host saw the job's batch
 * finish. */
enum {
  FHE_JOB_DOT = 0, /* multiply(a, b, rlk) then sum_1d_by_row: the encrypted
                      dot product of two slot vectors */
  FHE_JOB_POLY = 1 /* evaluate_polynomial(a) with the engine's coefficients */
};
typedef struct fhe_job {
  int engine;           /* index of the engine */
  int kind;             /* FHE_JOB_* */
  int level;            /* of the INTT-form, not rescaled inputs */
  const int64_t *a;     /* device duo [2][limbs][N] */
  const int64_t *b;     /* second duo (FHE_JOB_DOT) */
  int64_t *out;         /* device, room for [2][limbs][N] */
  double arrival_ms;
  double admitted_ms;   /* set by fhe_serve */
  double done_ms;       /* set by fhe_serve */
  int out_level, out_flags; /* set by fhe_serve */
} fhe_job;
typedef struct fhe_serve_engine {
  fhe_context *ctx;
  const int64_t *rlk;
  int rlk_moduli_id;
  fhe_rotation_key rotation;
  const double *coefficients; /* of the FHE_JOB_POLY jobs */
  int coefficient_count;
} fhe_serve_engine;
typedef struct fhe_serve_options {
  int max_batch;        /* jobs per batch */
  int max_in_flight;    /* batches per engine on the device at once */
  size_t reserve_bytes; /* free HBM kept back from admission */
} fhe_serve_options;
typedef struct fhe_serve_stats {
  double wall_ms;
  int batches;
  int max_batch_jobs;
  int deferred_admissions; /* admission checks that found too little HBM */
  size_t min_free_bytes;   /* least free HBM seen at an admission */
} fhe_serve_stats;
int fhe_serve(const fhe_serve_engine *engines, int engine_count,
              fhe_job *jobs, int job_count, const fhe_serve_options *options,
              fhe_serve_stats *stats);

/* Batches at uneven levels (N = 2^16).  The reference switches keys one
 * ciphertext at a time; these entry points take `count` ciphertexts, each at
 * its own moduli_id (and in its own device buffer), and run every stage of
 * the key switch as one launch over the whole batch: persistent kernels walk
 * per-shape work lists built on the first call of a shape and kept on the
 * device.  Every output equals the single-ciphertext call on that input.
 *
 * fhe_relinearize_uneven: CKKSCore::relinearize (src/ckks/core.cpp:713-741)
 * of NTT-form trios trios[c] [3][limbs(moduli_ids[c])][N] into INTT-form duos
 * outs[c] [2][limbs][N] (outs[c] must not alias trios[c]). */
int fhe_relinearize_uneven(fhe_context *ctx, int count,
                           const int *moduli_ids, const int64_t *const *trios,
                           const int64_t *key, int key_moduli_id,
                           int64_t *const *outs);

/* CKKSEngine::rotate(ct, FixedRotationKey*) (src/ckks/engine_rotate.cpp:
 * 55-76, CKKSCore::rotate, src/ckks/core.cpp:938-962) of INTT-form duos
 * ins[c] [2][limbs(moduli_ids[c])][N] at uneven levels with one fixed
 * rotation key made for `delta`; outs[c] may alias ins[c]. */
int fhe_rotate_uneven(fhe_context *ctx, int count, const int *moduli_ids,
                      const int64_t *const *ins, const int64_t *key,
                      int key_moduli_id, int delta, int64_t *const *outs);

/* CKKSEngine::rotate(ct, RotationKey*, delta) (src/ckks/engine_rotate.cpp:
 * 142-170) of INTT-form duos at uneven levels, ciphertext c by its own
 * deltas[c]: the RotationKey's fixed keys (keys[k] at key_moduli_ids[k],
 * made for key_deltas[k], in the key's order) rotate, key by key, the
 * ciphertexts whose normalized delta has that key's bit, all of them in one
 * launch sequence per key.  outs[c] may alias ins[c]; delta 0 copies. */
int fhe_rotate_rotation_key_uneven(fhe_context *ctx, int count,
                                   const int *moduli_ids,
                                   const int64_t *const *ins, int key_count,
                                   const int64_t *const *keys,
                                   const int *key_moduli_ids,
                                   const int *key_deltas, const int *deltas,
                                   int64_t *const *outs);

/* Launch schedules of the uneven-level kernels (no counterpart in the
 * reference).  With autotuning on, the first call of a batch shape times
 * each stage's schedule candidates - persistent grid caps of 1024, 2048 and
 * 4096 blocks or one block per work unit, with and without one contiguous
 * range of the work list per XCD - by running the whole key switch (one
 * pass of coordinate descent) and keeps the fastest; off, the default is
 * one block per unit with per-XCD ranges.  fhe_uneven_schedule reports the
 * schedule of a shape that has run (relin: the shape of
 * fhe_relinearize_uneven, else of the rotations): caps[s], xcds[s] for the
 * 8 stages (trio inverse pass 1 and 2, digits, extension, key product,
 * accumulator inverse pass 2, ModDown, automorphism). */
int fhe_context_uneven_autotune(fhe_context *ctx, int on);
int fhe_uneven_schedule(fhe_context *ctx, int relin, int count,
                        const int *moduli_ids, int key_moduli_id, int *caps,
                        int *xcds, int capacity);

/* HIP graph capture of a context's stream (no counterpart in the
 * reference): every call between fhe_graph_begin and fhe_graph_end is
 * recorded instead of run, and fhe_graph_launch replays the recording on the
 * context's stream with the same device buffers.  While capturing, calls
 * that would allocate, upload constants, grow a workspace or synchronize
 * fail with FHE_ERR_INVALID: run each call once before capturing it.  A
 * graph holds the addresses of the context's workspaces: while one exists,
 * a call that would grow a workspace fails too (destroy the graph first).
 * Destroy a context's graphs before the context. */
typedef struct fhe_graph fhe_graph;
int fhe_graph_begin(fhe_context *ctx);
int fhe_graph_end(fhe_context *ctx, fhe_graph **graph);
int fhe_graph_launch(fhe_context *ctx, fhe_graph *graph);
void fhe_graph_destroy(fhe_graph *graph);

/* Key switching of `count` INTT-form duos to the secret of a
 * key-switching key ([key_polys][key_limbs][N] at key_moduli_id): out =
 * (c0 + KS(c1)_0, KS(c1)_1) (CKKSCore::switch_key, src/ckks/core.cpp:
 * 1058-1083, used by mod_raise's dense/sparse switches,
 * src/ckks/engine_mod_raise.cpp:5-34).  out may alias in. */
int fhe_switch_key(fhe_context *ctx, int count, int moduli_id,
                   const int64_t *in, const int64_t *key, int key_moduli_id,
                   int64_t *out);

/* Conjugation of `count` INTT-form duos with a conjugation key
 * (CKKSEngine::conjugate, src/ckks/engine_conjugate.cpp:53-73;
 * CKKSCore::conjugate, src/ckks/core.cpp:1035-1056).  out may alias in. */
int fhe_conjugate(fhe_context *ctx, int count, int moduli_id,
                  const int64_t *in, const int64_t *key, int key_moduli_id,
                  int64_t *out);

/* Encoding of `count` slot vectors into plaintext limbs at `level`
 * (CKKSEngine::encode_unit, src/ckks/engine_encode.cpp:10-36, through
 * CKKSCore::encode_polynomial, src/ckks/core_encode.cpp:31-150): values is a
 * device array [count][slots] of doubles (complex_input = 0) or
 * [count][slots][2] (real, imaginary; complex_input = 1), slots = N / 2, the
 * caller's short vectors zero-padded as CKKSCore::pad does; out is
 * [count][limbs][N] INTT-form Montgomery limbs at
 * moduli_id = full_level - level.  Double arithmetic matches the reference's
 * operation for operation, so the limbs are bit-identical. */
int fhe_encode(fhe_context *ctx, const double *values, int complex_input,
               int count, int level, int64_t *out);
/* The per-level encoding constants of CKKSPrecomputed
 * (src/ckks/precomputed.cpp:138-215: get_ckks_scale, get_scale_before_tile,
 * get_log_scale_after_tile, get_r_mult_scale_after_tile over the level's
 * chain including the special primes) and whether the level is the top of a
 * one-more-level chain (src/ckks/core_encode.cpp:64-78); host only. */
int fhe_params_encode_constants(const fhe_params *params, int level,
                                long double *ckks_scale,
                                double *scale_before_tile,
                                int64_t *log_scale_after_tile,
                                int *top_of_one_more_level, int64_t *r_mult,
                                int capacity, int *count);

/* Decoding of `count` INTT-form plaintexts pt [count][limbs][N] at
 * moduli_id with the plaintexts' ckks_scale (CKKSEngine::decode_unit /
 * decode_complex_unit_plaintext, src/ckks/engine_decode.cpp:6-18, through
 * CKKSCore::convert_plaintext_to_double and decode_,
 * src/ckks/core_decode.cpp:10-124): the plaintexts are rescaled while
 * ckks_scale / q stays above the base scale, the limbs are lifted to centered
 * integers by CRT, and the inverse canonical embedding gives
 * out [count][slots] real parts (complex_output = 0) or [count][slots][2]
 * (complex_output = 1), device memory.  The reference's errors: fewer than
 * two limbs, or a scale above the limbs' modulus product.  pt is not
 * modified. */
int fhe_decode(fhe_context *ctx, const int64_t *pt, int count, int moduli_id,
               long double ckks_scale, int complex_output, double *out);
/* The moduli_id and scale the decoder reaches after its rescales (host
 * only; the same loop as fhe_decode). */
int fhe_decoded_scale(const fhe_context *ctx, int moduli_id,
                      long double ckks_scale, int *decoded_moduli_id,
                      long double *decoded_scale);

/* Multiplication of `count` INTT-form duos at moduli_id by value * i, the
 * imaginary unit times an integer (CKKSEngine::multiply_imaginary,
 * src/ckks/engine_multiply.cpp:468-491): the ciphertext and the plaintext
 * value * X^(N/2) (CKKSCore::encode_imaginary, src/ckks/core_encode.cpp:
 * 190-210) are taken to NTT form, multiplied and brought back.  The scale is
 * unchanged.  out may alias in. */
int fhe_multiply_imaginary(fhe_context *ctx, int count, int moduli_id,
                           const int64_t *in, int64_t value, int64_t *out);

/* Modulus raising of `count` INTT-form duos at level 0 for bootstrapping
 * (CKKSEngine::mod_raise_, src/ckks/engine_mod_raise.cpp:5-34), on the
 * not-rescaled level-0 form that CKKSEngine::mod_raise passes after
 * level_down(ct, 0) (moduli_id full_level, two limbs): rescale to the base
 * prime, switch to the sparse secret with to_sparse (a key at
 * sparse_moduli_id), lift the
 * signed residues onto every prime of the maximum key-switching level
 * (CKKSCore::make_signed_tile, src/ckks/core.cpp:1085-1110), switch back
 * with to_dense (at dense_moduli_id) and multiply by the integer ratio
 * (CKKSCore::get_mod_raise_ratio, src/ckks/core.cpp:1112-1119).  out:
 * [count][2][limbs][N] at the moduli_id of the maximum key-switching level;
 * its scale is scale^2 * ratio. */
int fhe_mod_raise(fhe_context *ctx, int count, const int64_t *in,
                  const int64_t *to_sparse, int sparse_moduli_id,
                  const int64_t *to_dense, int dense_moduli_id,
                  int64_t *out);

/* Pipeline choices of a context (no counterpart in the reference: they
 * select among equivalent device implementations, all bit-exact), for tuning
 * and A/B measurement.  Names and defaults:
 *   "lazy_key_switch" 1    lazy transforms where results are canonicalized
 *                          later;
 *   "staged_modup" 1       ModUp through the Garner digits;
 *   "fuse_modup" 0, "fuse_tensor" 0   one-kernel variants (slower);
 *   "own_limbs" 1          partitions' own extended limbs from the NTT-form
 *                          input;
 *   "split_variants" 1     one NTT launch per arithmetic variant;
 *   "ks_lazy_accumulate" 1 small-prime key products summed without
 *                          intermediate reductions, large-prime sums kept
 *                          below 2q with one conditional subtraction;
 *   "ks_wide_accumulate" 0 small primes' key products summed as unreduced
 *                          96-bit values with one Montgomery reduction per
 *                          accumulator (fewer VALU instructions, measured
 *                          slower);
 *   "overlap_trio" 1       N = 2^16 with split_variants: the INTT of d0 and
 *                          d1 in fhe_hmult_relin_rescale runs on a second
 *                          non-blocking stream of the context (default
 *                          priority), joined by an event before ModDown;
 *                          forked once the trio is written (1), or once
 *                          d2's INTT is done (2: beside the key inner
 *                          product, about 1% more throughput, that kernel's
 *                          launches then sharing the device); 0: one stream;
 *   "ks_occupancy" 2       waves per SIMD of the key inner product (3 trades
 *                          registers for occupancy; slower);
 *   "ks_prefetch" 0        the key inner product loads the next partition's
 *                          tile ahead (slower).
 * Synchronizes the context's stream. */
int fhe_context_set_option(fhe_context *ctx, const char *name, int value);

/* Calibration (no counterpart in the reference): the device's rate of
 * modular products, measured by a kernel of independent product chains at
 * full occupancy; form 0 = the Shoup product of the lazy transforms, 1 =
 * the Montgomery product of the exact ones.  The hot path is bound by this
 * rate rather than by HBM; benchmarks divide their algorithmic products per
 * second by it.  Synchronizes the context's stream. */
int fhe_modmul_peak(fhe_context *ctx, int form, double *per_second);

/* Stage timing (no counterpart in the reference, which has no device
 * stages): while enabled, every pipeline stage that a call launches on the
 * context's stream is bracketed by HIP events; fhe_profile_count / _entry
 * synchronize the stream and report each stage's name, total milliseconds
 * and launch count since the last fhe_profile_enable(ctx, 1). */
int fhe_profile_enable(fhe_context *ctx, int on);
int fhe_profile_count(fhe_context *ctx, int *count);
int fhe_profile_entry(fhe_context *ctx, int index, char *name,
                      int name_capacity, double *total_ms, int64_t *launches);

/* Wire format of device data for storage and exchange.  The reference has
 * none (Data::write, src/data/data.h:66, is declared and never defined);
 * this one records the prime chain (a fingerprint of the primes and log2 N),
 * the shape and form (polys, moduli_id, limbs, level, data flags), the
 * long double scale and a payload checksum ahead of the limbs in the
 * reference's [poly][limb][N] order.  fhe_wire_write copies the device
 * limbs into a host buffer of fhe_wire_size bytes; fhe_wire_read checks the
 * header and the checksum against this context and copies the limbs to a
 * device buffer. */
enum { FHE_WIRE_PLAINTEXT = 0, FHE_WIRE_CIPHERTEXT = 1, FHE_WIRE_KEY = 2 };
typedef struct fhe_wire_header {
  int kind, polys, moduli_id, limbs, level, flags;
  long double scale;
  size_t payload_bytes;
} fhe_wire_header;
int fhe_wire_size(const fhe_context *ctx, int polys, int limbs,
                  size_t *bytes);
int fhe_wire_write(fhe_context *ctx, int kind, const int64_t *data, int polys,
                   int moduli_id, int limbs, int level, int flags,
                   long double scale, void *buffer, size_t capacity,
                   size_t *written);
int fhe_wire_read_header(const fhe_context *ctx, const void *buffer,
                         size_t size, fhe_wire_header *header);
int fhe_wire_read(fhe_context *ctx, const void *buffer, size_t size,
                  int64_t *data, size_t capacity);

#ifdef __cplusplus
}
#endif

#endif /* FHE_AMD_H */
