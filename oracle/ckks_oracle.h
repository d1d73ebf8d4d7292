/*
 * ckks_oracle.h -- CPU restatement of the reference's CKKS hot-path
 * arithmetic, used ONLY as a test checker.
 *
 * This library is test infrastructure.  Only tests/, the smoke check in
 * __graft_entry__.py and bench.py's cpu_baseline leg may load it; the GPU
 * product (fhe_amd/) never links or calls it.
 *
 * Every function states the reference file:line it follows.  Limb arrays are
 * int64 in the reference's order [polynomial][limb][coefficient]; limb i of an
 * object at moduli_id m is a residue modulo prime m + i of the full chain
 * (src/ckks/precomputed.cpp:227-230), special primes follow the Q chain.
 */
#ifndef CKKS_ORACLE_H
#define CKKS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_ctx oracle_ctx;

/* scalar Montgomery unit (src/montgomery/unit.h:13-31) */
int64_t oracle_mont_mult(int64_t a, int64_t b, uint64_t q, uint64_t k);
int64_t oracle_mont_exit(int64_t x, uint64_t q, uint64_t k);
uint64_t oracle_calculate_k(uint64_t q);
int64_t oracle_find_root(int64_t q, int coeff_count);

/* raw per-limb transforms with caller twiddles (src/operation/ntt.cpp) */
void oracle_ntt_raw(int64_t *x, const int64_t *root, uint64_t q, uint64_t k,
                    int log_n);
void oracle_intt_raw(int64_t *x, const int64_t *root_inverse, uint64_t q,
                     uint64_t k, int log_n);

/* context: a prime chain q[0..prime_count) whose last `special_count`
 * entries are the key-switching special primes */
oracle_ctx *oracle_ctx_create(int log_n, const uint64_t *primes,
                              int prime_count, int special_count);
void oracle_ctx_destroy(oracle_ctx *ctx);
int oracle_ctx_chain_count(const oracle_ctx *ctx);

/* out = {k, r_squared, coeff_count_inverse}; root / root_inverse: N-1 each */
void oracle_prime_constants(const oracle_ctx *ctx, int prime_index,
                            int64_t *out, int64_t *root,
                            int64_t *root_inverse);

void oracle_ntt_limb(const oracle_ctx *ctx, int prime_index, int64_t *x);
void oracle_intt_limb(const oracle_ctx *ctx, int prime_index, int64_t *x);

/* domain changes over `limbs` limbs starting at prime `moduli_id` */
void oracle_to_ntt(const oracle_ctx *ctx, int moduli_id, int limbs,
                   int64_t *x);
void oracle_to_intt(const oracle_ctx *ctx, int moduli_id, int limbs,
                    int64_t *x);

/* elementwise Montgomery product of `limbs` limbs starting at moduli_id */
void oracle_mont_mult_limbs(const oracle_ctx *ctx, int moduli_id, int limbs,
                            const int64_t *a, const int64_t *b, int64_t *out);

/* NTT-form tensor of two duos -> trio */
void oracle_tensor(const oracle_ctx *ctx, int moduli_id, int limbs,
                   const int64_t *a, const int64_t *b, int64_t *trio);

/* hybrid key switching of one INTT-form polynomial x (limbs = chain count at
 * moduli_id) against a key at key_moduli_id -> two INTT-form polynomials */
void oracle_key_switch(const oracle_ctx *ctx, int moduli_id, const int64_t *x,
                       const int64_t *key, int key_moduli_id, int64_t *out);

/* relinearize an NTT-form trio -> INTT-form duo */
void oracle_relinearize(const oracle_ctx *ctx, int moduli_id,
                        const int64_t *trio, const int64_t *key,
                        int key_moduli_id, int64_t *duo);

/* rescale `polys` INTT-form polynomials at moduli_id: drops the front limb */
void oracle_rescale(const oracle_ctx *ctx, int moduli_id, int polys,
                    const int64_t *x, int64_t *y);

/* one HMult+relin+rescale unit: a, b rescaled INTT duos at moduli_id ->
 * out duo at moduli_id + 1 */
void oracle_hmult_relin_rescale(const oracle_ctx *ctx, int moduli_id,
                                const int64_t *a, const int64_t *b,
                                const int64_t *key, int key_moduli_id,
                                int64_t *out);

/* coefficient-domain automorphism X -> X^(5^-delta) (delta = -1:
 * conjugation) of polys x limbs INTT-form limbs */
void oracle_automorphism(const oracle_ctx *ctx, int moduli_id, int limbs,
                         int polys, int delta, const int64_t *x, int64_t *y);

/* rotation of an INTT-form duo with a fixed rotation key made for `delta`,
 * and conjugation with a conjugation key; outputs INTT-form duos */
void oracle_rotate(const oracle_ctx *ctx, int moduli_id, const int64_t *ct,
                   const int64_t *key, int key_moduli_id, int delta,
                   int64_t *out);
void oracle_conjugate(const oracle_ctx *ctx, int moduli_id, const int64_t *ct,
                      const int64_t *key, int key_moduli_id, int64_t *out);
void oracle_switch_key(const oracle_ctx *ctx, int moduli_id, const int64_t *ct,
                       const int64_t *key, int key_moduli_id, int64_t *out);

/* public-key encryption with explicit samples u, e0, e1 (N values each)
 * and duo decryption to plaintext limbs */
void oracle_encrypt_pk(const oracle_ctx *ctx, int moduli_id, const int64_t *pt,
                       const int64_t *pk, int pk_moduli_id, const int64_t *u,
                       const int64_t *e0, const int64_t *e1, int64_t *out);
void oracle_decrypt(const oracle_ctx *ctx, int moduli_id, const int64_t *ct,
                    const int64_t *sk, int sk_moduli_id, int64_t *out);

/* CKKS encoding (src/ckks/core_encode.cpp:11-150): the slot vector
 * (n/2 doubles, or n/2 {re, im} pairs when complex_input) at a level whose
 * constants are given: scale_before_tile, and r_mult (the level's
 * r_mult_scale_after_tile, one per encoded limb).  out [out_limbs][n]:
 * limbs [0, first) zero, limb first + e on prime enc_moduli_id + e. */
void oracle_encode(const oracle_ctx *ctx, const double *values,
                   int complex_input, double scale_before_tile, int first,
                   int enc_moduli_id, int out_limbs, const int64_t *r_mult,
                   int64_t *out);
/* the slot -> codec index map of the encoder (src/ckks/encode.cpp:24-150,
 * src/ckks/precomputed.cpp:50-58), codec = n */
void oracle_encode_pre_permutation(int codec, uint64_t *pre);

/* CKKS decoding (src/ckks/core_decode.cpp:10-124) of one plaintext pt
 * [chain_count - moduli_id][n] with its long double ckks_scale and the base
 * scale 2^scale_bits: the rescale loop, the CRT lift, the inverse canonical
 * embedding.  out: n/2 doubles, or n/2 {re, im} pairs when complex_output.
 * Returns 0, or -1 for the reference's errors (one limb, scale overflow). */
int oracle_decode(const oracle_ctx *ctx, const int64_t *pt, int moduli_id,
                  long double ckks_scale, double base_scale, int complex_output,
                  double *out);

/* its CRT step alone: limbs at moduli_id -> n signed values as doubles */
void oracle_decode_crt(const oracle_ctx *ctx, const int64_t *pt, int moduli_id,
                       double *out);

/* one limb of a per-coefficient primitive (codes of FHE_EW_*) */
void oracle_elementwise_limb(int op, const int64_t *x1, const int64_t *x2,
                             int64_t *y, int n, uint64_t q, uint64_t k,
                             int64_t c);

/* partition layout for a ciphertext at moduli_id against a key at
 * key_moduli_id: returns partition count; starts/sizes/key_index filled */
int oracle_partitions(const oracle_ctx *ctx, int moduli_id, int key_moduli_id,
                      int *starts, int *sizes, int *key_partition_index);

#ifdef __cplusplus
}
#endif

#endif
