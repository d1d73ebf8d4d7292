/* Host sanitizer run of the oracle (test infrastructure only): built with
 * -fsanitize=address,undefined -fno-sanitize-recover=all together with
 * ckks_oracle.c by tests/test_oracle_sanitized.py, it drives every oracle
 * entry point on a small chain (N = 2^10, 6 chain primes, 2 special primes,
 * random limbs and keys from a fixed seed), so that an out-of-bounds access,
 * a leak or undefined behaviour in the checker ends the run with a nonzero
 * status.  It also checks what needs no reference data: to_intt(to_ntt(x))
 * = x, and that rescale outputs are canonical.  Prints "ok" and returns 0
 * when clean. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ckks_oracle.h"

enum { kLogN = 10, kN = 1 << kLogN, kChain = 6, kAlpha = 2 };
enum { kPrimes = kChain + kAlpha };

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint64_t next_u64(void) {
  /* xorshift64* */
  rng_state ^= rng_state >> 12;
  rng_state ^= rng_state << 25;
  rng_state ^= rng_state >> 27;
  return rng_state * 0x2545F4914F6CDD1Dull;
}

static uint64_t mul_mod(uint64_t a, uint64_t b, uint64_t m) {
  return (uint64_t)((unsigned __int128)a * b % m);
}
static uint64_t pow_mod(uint64_t a, uint64_t e, uint64_t m) {
  uint64_t r = 1 % m;
  while (e) {
    if (e & 1)
      r = mul_mod(r, a, m);
    a = mul_mod(a, a, m);
    e >>= 1;
  }
  return r;
}
/* deterministic Miller-Rabin for 64-bit n */
static int is_prime(uint64_t n) {
  static const uint64_t bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  if (n < 2)
    return 0;
  uint64_t d = n - 1;
  int s = 0;
  while ((d & 1) == 0) {
    d >>= 1;
    s++;
  }
  for (size_t i = 0; i < sizeof(bases) / sizeof(bases[0]); i++) {
    const uint64_t a = bases[i] % n;
    if (a == 0)
      continue;
    uint64_t x = pow_mod(a, d, n);
    if (x == 1 || x == n - 1)
      continue;
    int composite = 1;
    for (int r = 1; r < s; r++) {
      x = mul_mod(x, x, n);
      if (x == n - 1) {
        composite = 0;
        break;
      }
    }
    if (composite)
      return 0;
  }
  return 1;
}

/* primes = 1 mod 2N below 2^bits, descending */
static void find_primes(int bits, int count, uint64_t *out) {
  const uint64_t step = 2 * (uint64_t)kN;
  uint64_t c = ((1ull << bits) / step) * step + 1;
  int found = 0;
  while (found < count) {
    c -= step;
    if (is_prime(c))
      out[found++] = c;
  }
}

static int64_t *limbs_alloc(int polys, int limbs) {
  return calloc((size_t)polys * limbs * kN, sizeof(int64_t));
}

/* random residues: limb l of every polynomial below primes[mid + l] */
static void fill(int64_t *x, int polys, int limbs, const uint64_t *primes,
                 int mid) {
  for (int p = 0; p < polys; p++)
    for (int l = 0; l < limbs; l++)
      for (int t = 0; t < kN; t++)
        x[((size_t)p * limbs + l) * kN + t] =
            (int64_t)(next_u64() % primes[mid + l]);
}

static int fail(const char *what) {
  fprintf(stderr, "sanitize_check: %s\n", what);
  return 1;
}

int main(void) {
  uint64_t primes[kPrimes];
  find_primes(50, kChain, primes);
  find_primes(52, kAlpha, primes + kChain);
  oracle_ctx *c = oracle_ctx_create(kLogN, primes, kPrimes, kAlpha);
  if (oracle_ctx_chain_count(c) != kChain)
    return fail("chain count");

  /* per-prime constants and the limb transforms */
  int64_t cst[3];
  int64_t *root = malloc(sizeof(int64_t) * (kN - 1));
  int64_t *root_inv = malloc(sizeof(int64_t) * (kN - 1));
  for (int i = 0; i < kPrimes; i++)
    oracle_prime_constants(c, i, cst, root, root_inv);
  int64_t *x = limbs_alloc(1, kChain), *x0 = limbs_alloc(1, kChain);
  fill(x, 1, kChain, primes, 0);
  memcpy(x0, x, sizeof(int64_t) * kChain * kN);
  oracle_to_ntt(c, 0, kChain, x);
  oracle_to_intt(c, 0, kChain, x);
  for (int l = 0; l < kChain; l++)
    for (int t = 0; t < kN; t++) {
      const int64_t q = (int64_t)primes[l];
      int64_t v = x[(size_t)l * kN + t] % q;
      if (v < 0)
        v += q;
      if (v != x0[(size_t)l * kN + t])
        return fail("to_intt(to_ntt(x)) != x");
    }
  for (int i = 0; i < kPrimes; i++) {
    oracle_ntt_limb(c, i, x);
    oracle_intt_limb(c, i, x);
  }

  /* the chain operations at moduli_id 1 against keys at moduli_id 0 */
  const int mid = 1, limbs = kChain - mid, key_mid = 0;
  const int key_chain = kChain - key_mid + kAlpha;
  int starts[64], sizes[64], kidx[64];
  const int key_parts = oracle_partitions(c, key_mid, key_mid, starts, sizes,
                                          kidx);
  const int parts = oracle_partitions(c, mid, key_mid, starts, sizes, kidx);
  if (parts < 1 || parts > key_parts)
    return fail("partitions");
  int64_t *key = calloc((size_t)key_parts * 2 * key_chain * kN,
                        sizeof(int64_t));
  for (size_t i = 0; i < (size_t)key_parts * 2 * key_chain * kN; i++)
    key[i] = (int64_t)(next_u64() % primes[0]);
  int64_t *a = limbs_alloc(2, limbs), *b = limbs_alloc(2, limbs);
  int64_t *trio = limbs_alloc(3, limbs), *duo = limbs_alloc(2, limbs);
  int64_t *out = limbs_alloc(2, limbs), *ks = limbs_alloc(2, limbs);
  int64_t *resc = limbs_alloc(2, limbs - 1);
  fill(a, 2, limbs, primes, mid);
  fill(b, 2, limbs, primes, mid);
  oracle_mont_mult_limbs(c, mid, limbs, a, b, out);
  oracle_tensor(c, mid, limbs, a, b, trio);
  oracle_relinearize(c, mid, trio, key, key_mid, duo);
  oracle_key_switch(c, mid, a + (size_t)limbs * kN, key, key_mid, ks);
  oracle_rescale(c, mid, 2, duo, resc);
  for (int l = 0; l < limbs - 1; l++)
    for (int t = 0; t < 2 * kN; t++) {
      const int64_t v = resc[((size_t)(t / kN) * (limbs - 1) + l) * kN +
                             t % kN];
      if (v < 0 || (uint64_t)v >= primes[mid + 1 + l])
        return fail("rescale output not canonical");
    }
  oracle_hmult_relin_rescale(c, mid, a, b, key, key_mid, out);
  oracle_automorphism(c, mid, limbs, 2, 5, a, out);
  oracle_automorphism(c, mid, limbs, 2, -1, a, out);
  oracle_rotate(c, mid, a, key, key_mid, 3, out);
  oracle_conjugate(c, mid, a, key, key_mid, out);
  oracle_switch_key(c, mid, a, key, key_mid, out);

  /* encryption with explicit samples and decryption */
  int64_t *pt = limbs_alloc(1, limbs), *pk = limbs_alloc(2, kChain);
  int64_t *sk = limbs_alloc(1, kChain), *dec = limbs_alloc(1, limbs);
  int64_t *u = calloc(kN, sizeof(int64_t)), *e0 = calloc(kN, sizeof(int64_t)),
          *e1 = calloc(kN, sizeof(int64_t));
  fill(pt, 1, limbs, primes, mid);
  fill(pk, 2, kChain, primes, 0);
  for (int t = 0; t < kN; t++) {
    u[t] = (int64_t)(next_u64() % 3) - 1;
    e0[t] = (int64_t)(next_u64() % 7) - 3;
    e1[t] = (int64_t)(next_u64() % 7) - 3;
    sk[t] = (int64_t)(next_u64() % 3) - 1;
  }
  for (int l = 1; l < kChain; l++)
    memcpy(sk + (size_t)l * kN, sk, sizeof(int64_t) * kN);
  oracle_encrypt_pk(c, mid, pt, pk, 0, u, e0, e1, out);
  oracle_decrypt(c, mid, out, sk, 0, dec);

  /* encoding and decoding */
  double *values = malloc(sizeof(double) * kN);
  for (int i = 0; i < kN; i++)
    values[i] = (double)(next_u64() % 2001) / 1000.0 - 1.0;
  uint64_t *pre = malloc(sizeof(uint64_t) * kN); /* codec = N entries */
  oracle_encode_pre_permutation(kN, pre);
  int64_t *r_mult = malloc(sizeof(int64_t) * limbs);
  for (int l = 0; l < limbs; l++)
    r_mult[l] = (int64_t)(next_u64() % primes[mid + l]);
  int64_t *enc = limbs_alloc(1, limbs);
  oracle_encode(c, values, 0, 1099511627776.0, 0, mid, limbs, r_mult, enc);
  oracle_encode(c, values, 1, 1099511627776.0, 0, mid, limbs, r_mult, enc);
  double *decoded = malloc(sizeof(double) * kN);
  double *crt = malloc(sizeof(double) * kN);
  oracle_decode_crt(c, enc, mid, crt);
  (void)oracle_decode(c, enc, mid, 1099511627776.0L, 1099511627776.0, 0,
                      decoded);
  (void)oracle_decode(c, enc, mid, 1099511627776.0L, 1099511627776.0, 1,
                      decoded);

  /* every elementwise code on one limb */
  const uint64_t q = primes[mid];
  int64_t ck[3];
  oracle_prime_constants(c, mid, ck, root, root_inv);
  for (int op = 0; op <= 21; op++) {
    memcpy(out, a, sizeof(int64_t) * kN);
    oracle_elementwise_limb(op, a, b, out, kN, q, (uint64_t)ck[0], 12345);
  }
  for (int t = 0; t < 8; t++) {
    const int64_t v = (int64_t)(next_u64() % q);
    (void)oracle_mont_exit(oracle_mont_mult(v, v, q, (uint64_t)ck[0]), q,
                           (uint64_t)ck[0]);
  }

  free(crt);
  free(decoded);
  free(enc);
  free(r_mult);
  free(pre);
  free(values);
  free(e1);
  free(e0);
  free(u);
  free(dec);
  free(sk);
  free(pk);
  free(pt);
  free(resc);
  free(ks);
  free(out);
  free(duo);
  free(trio);
  free(b);
  free(a);
  free(key);
  free(x0);
  free(x);
  free(root_inv);
  free(root);
  oracle_ctx_destroy(c);
  printf("ok\n");
  return 0;
}
