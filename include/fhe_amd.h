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

/* Conjugation of `count` INTT-form duos with a conjugation key
 * (CKKSEngine::conjugate, src/ckks/engine_conjugate.cpp:53-73;
 * CKKSCore::conjugate, src/ckks/core.cpp:1035-1056).  out may alias in. */
int fhe_conjugate(fhe_context *ctx, int count, int moduli_id,
                  const int64_t *in, const int64_t *key, int key_moduli_id,
                  int64_t *out);

/* Stage timing (no counterpart in the reference, which has no device
 * stages): while enabled, every pipeline stage that a call launches on the
 * context's stream is bracketed by HIP events; fhe_profile_count / _entry
 * synchronize the stream and report each stage's name, total milliseconds
 * and launch count since the last fhe_profile_enable(ctx, 1). */
int fhe_profile_enable(fhe_context *ctx, int on);
int fhe_profile_count(fhe_context *ctx, int *count);
int fhe_profile_entry(fhe_context *ctx, int index, char *name,
                      int name_capacity, double *total_ms, int64_t *launches);

#ifdef __cplusplus
}
#endif

#endif /* FHE_AMD_H */
