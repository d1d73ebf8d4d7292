/*
 * ckks_oracle.c -- CPU restatement of the reference's CKKS hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the GPU product is compared
 * against; the product never links it.  Only tests/, the smoke check and
 * bench.py's cpu_baseline leg may load it.
 *
 * It restates, in plain C, the scalar semantics and the exact operation order
 * of the reference (paths relative to the reference root):
 *   Montgomery unit / compounds   src/montgomery/unit.h:13-31,
 *                                 src/montgomery/unit.cpp, compound.cpp
 *   NTT / INTT                    src/operation/ntt.cpp:46-134, :177-272
 *   per-prime constants           src/ckks/derived_parameters.cpp:19-50,
 *                                 src/arithmetic/arithmetic.cpp:10-167
 *   chain-level constants         src/ckks/precomputed.cpp:33-137, :325-382
 *   partitions                    src/ckks/parameters.cpp:325-360
 *   tensor                        src/ckks/core.cpp:284-314
 *   to_ntt / to_intt              src/ckks/core_ntt.cpp:91-150
 *   MRS / ModUp / ModDown / KS    src/ckks/core_key_switch.cpp:21-160, :293-358
 *   relinearize                   src/ckks/core.cpp:713-741
 *   rescale                       src/ckks/core_rescale.cpp:5-25
 *
 * Values are int64 with the reference's two's-complement behaviour: the
 * Montgomery product sign-extends its operands to 128 bits exactly like
 * static_cast<uint128_t>(int64_t) does, so negative intermediates in ModDown
 * and rescale reproduce bit for bit.
 */
#define _GNU_SOURCE /* sincos */
#include "ckks_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;
typedef __int128 i128;

#define NBITS 62
#define FB_MASK ((UINT64_C(1) << NBITS) - 1)
#define R_VALUE (INT64_C(1) << NBITS)

/* ---------------------------------------------------------------- scalars */

/* src/montgomery/unit.h:13 */
static inline int64_t guard(int64_t x, int64_t q) { return (x < q) ? x : x - q; }

/* src/montgomery/unit.h:15-23 */
int64_t oracle_mont_mult(int64_t a, int64_t b, uint64_t q, uint64_t k) {
  const u128 x = (u128)(i128)a * (u128)(i128)b;
  const uint64_t s = (uint64_t)(x * k) & FB_MASK;
  const u128 sq = (u128)s * (u128)q;
  return (int64_t)(uint64_t)((x + sq) >> NBITS);
}

/* src/montgomery/unit.h:25-31 */
int64_t oracle_mont_exit(int64_t x, uint64_t q, uint64_t k) {
  const uint64_t s = ((uint64_t)x * k) & FB_MASK;
  const u128 sq = (u128)s * (u128)q;
  return (int64_t)(uint64_t)(((u128)(i128)x + sq) >> NBITS);
}

#define MONT oracle_mont_mult

/* src/arithmetic/arithmetic.cpp:10-31 (extended Euclid, 0 if not invertible) */
static int64_t modulus_invert(int64_t x, int64_t modulus) {
  int64_t a = x, b = modulus, ra = 1, rb = 0;
  while (b > 0) {
    int64_t quotient = a / b;
    int64_t t = a;
    a = b;
    b = t - quotient * b;
    t = ra;
    ra = rb;
    rb = t - quotient * rb;
  }
  if (a != 1)
    return 0;
  return ra >= 0 ? ra : ra + modulus;
}

/* src/arithmetic/arithmetic.cpp:33-38 */
static int64_t modulus_multiply(int64_t a, int64_t b, int64_t m) {
  i128 p = (i128)a * (i128)b;
  p %= (i128)m;
  return (int64_t)p;
}

/* src/arithmetic/arithmetic.cpp:40-52 */
static int64_t modulus_power(int64_t base, int64_t e, int64_t m) {
  int64_t cur = 1;
  for (; e > 0; e >>= 1) {
    if (e & 1)
      cur = modulus_multiply(cur, base, m);
    base = modulus_multiply(base, base, m);
  }
  return cur;
}

/* src/arithmetic/arithmetic.cpp:98-113: first base whose (q-1)/2N power has
 * order exactly 2N */
int64_t oracle_find_root(int64_t q, int coeff_count) {
  const int64_t unit_power = (q - 1) / (2 * (int64_t)coeff_count);
  int64_t e = 0;
  for (int64_t base = 2; base < q; base++) {
    e = modulus_power(base, unit_power, q);
    if (modulus_power(e, coeff_count, q) != 1)
      break;
  }
  return e;
}

/* src/arithmetic/arithmetic.cpp:159-167: k = (R * R^-1 - 1) / q */
uint64_t oracle_calculate_k(uint64_t q) {
  const int64_t rinv = modulus_invert(R_VALUE, (int64_t)q);
  i128 k = (i128)R_VALUE * (i128)rinv - 1;
  k /= (i128)q;
  return (uint64_t)k;
}

static int reverse_bits(int x, int bits) {
  int y = 0;
  for (int i = 0; i < bits; i++)
    y |= ((x >> i) & 1) << (bits - 1 - i);
  return y;
}

/* ------------------------------------------------------------------- NTT */

/* src/operation/ntt.cpp:46-134.  Stage s has 2^s groups of span N/2^(s+1);
 * the reference walks the twiddle array sequentially, i.e. group g of stage
 * s uses root[2^s - 1 + g]. */
void oracle_ntt_raw(int64_t *y, const int64_t *root, uint64_t q, uint64_t k,
                    int log_n) {
  const int n = 1 << log_n;
  const int64_t q2 = (int64_t)(q << 1);
  for (int s = 0; s < log_n; s++) {
    const int groups = 1 << s;
    const int span = n >> (s + 1);
    for (int g = 0; g < groups; g++) {
      const int64_t w = root[groups - 1 + g];
      int64_t *lo = y + 2 * g * span, *hi = lo + span;
      for (int i = 0; i < span; i++) {
        const int64_t a = lo[i];
        const int64_t p = MONT(w, hi[i], q, k);
        lo[i] = guard(a + p, q2);
        hi[i] = guard(a + q2 - p, q2);
      }
    }
  }
}

/* src/operation/ntt.cpp:177-272.  Stage t (span 2^t, N/2^(t+1) groups) uses
 * root_inverse[N - N/2^t + g] (the backward paint order of
 * src/ckks/precomputed.cpp:15-31). */
void oracle_intt_raw(int64_t *y, const int64_t *root_inverse, uint64_t q,
                     uint64_t k, int log_n) {
  const int n = 1 << log_n;
  const int64_t q2 = (int64_t)(q << 1);
  int offset = 0;
  for (int t = 0; t < log_n; t++) {
    const int span = 1 << t;
    const int groups = n >> (t + 1);
    for (int g = 0; g < groups; g++) {
      const int64_t w = root_inverse[offset + g];
      int64_t *lo = y + 2 * g * span, *hi = lo + span;
      for (int i = 0; i < span; i++) {
        const int64_t a = lo[i], b = hi[i];
        lo[i] = guard(a + b, q2);
        hi[i] = MONT(w, guard(a + q2 - b, q2), q, k);
      }
    }
    offset += groups;
  }
}

/* --------------------------------------------------------------- context */

struct oracle_ctx {
  int log_n, n;
  int prime_count, chain_count, alpha;
  uint64_t *q, *k, *q2, *qh;
  int64_t *r2, *ninv;
  int64_t *root, *root_inv;   /* [prime][n-1] */
  int64_t *p_product_mult_r;  /* [chain_count] */
  int64_t *p_inverse_mult_r;  /* [alpha][prime_count]: sp-th special from the
                                 end, entry per absolute prime index */
  int64_t *rescale_scale;     /* [prime_count][prime_count]: [from][to] */
};

int oracle_ctx_chain_count(const oracle_ctx *ctx) { return ctx->chain_count; }

oracle_ctx *oracle_ctx_create(int log_n, const uint64_t *primes,
                              int prime_count, int special_count) {
  oracle_ctx *c = (oracle_ctx *)calloc(1, sizeof(oracle_ctx));
  const int n = 1 << log_n;
  c->log_n = log_n;
  c->n = n;
  c->prime_count = prime_count;
  c->alpha = special_count;
  c->chain_count = prime_count - special_count;
  c->q = malloc(sizeof(uint64_t) * prime_count);
  c->k = malloc(sizeof(uint64_t) * prime_count);
  c->q2 = malloc(sizeof(uint64_t) * prime_count);
  c->qh = malloc(sizeof(uint64_t) * prime_count);
  c->r2 = malloc(sizeof(int64_t) * prime_count);
  c->ninv = malloc(sizeof(int64_t) * prime_count);
  c->root = malloc(sizeof(int64_t) * (size_t)prime_count * (n - 1));
  c->root_inv = malloc(sizeof(int64_t) * (size_t)prime_count * (n - 1));
  int64_t *power = malloc(sizeof(int64_t) * n);
  int64_t *shuffled = malloc(sizeof(int64_t) * n);

  for (int i = 0; i < prime_count; i++) {
    /* src/ckks/derived_parameters.cpp:19-50 */
    const uint64_t q = primes[i];
    c->q[i] = q;
    c->q2[i] = q << 1;
    c->qh[i] = q >> 1;
    c->r2[i] = modulus_multiply(R_VALUE, R_VALUE, (int64_t)q);
    c->k[i] = oracle_calculate_k(q);
    c->ninv[i] = modulus_multiply(modulus_invert(n, (int64_t)q), R_VALUE,
                                  (int64_t)q);
    const int64_t psi = oracle_find_root((int64_t)q, n);
    for (int pass = 0; pass < 2; pass++) {
      const int64_t base = pass == 0 ? psi : modulus_invert(psi, (int64_t)q);
      power[0] = 1;
      for (int j = 1; j < n; j++)
        power[j] = modulus_multiply(power[j - 1], base, (int64_t)q);
      for (int j = 0; j < n; j++)
        shuffled[reverse_bits(j, log_n)] = power[j];
      int64_t *dst = (pass == 0 ? c->root : c->root_inv) + (size_t)i * (n - 1);
      if (pass == 0) {
        for (int j = 1; j < n; j++)
          dst[j - 1] = shuffled[j];
      } else {
        /* backward paint: groups N/2.., N/4.., ..., 1 */
        int o = 0;
        for (int lg = log_n - 1; lg >= 0; lg--) {
          const int count = 1 << lg;
          for (int g = 0; g < count; g++)
            dst[o++] = shuffled[count + g];
        }
      }
      for (int j = 0; j < n - 1; j++)
        dst[j] = MONT(dst[j], c->r2[i], q, c->k[i]);
    }
  }
  free(power);
  free(shuffled);

  /* src/ckks/precomputed.cpp:97-107 */
  c->p_product_mult_r = malloc(sizeof(int64_t) * c->chain_count);
  for (int i = 0; i < c->chain_count; i++) {
    int64_t product = R_VALUE;
    for (int p = c->chain_count; p < prime_count; p++)
      product = modulus_multiply(product, (int64_t)c->q[p], (int64_t)c->q[i]);
    c->p_product_mult_r[i] = product;
  }
  /* src/ckks/precomputed.cpp:109-121 */
  c->p_inverse_mult_r = calloc((size_t)special_count * prime_count,
                               sizeof(int64_t));
  for (int sp = 0; sp < special_count; sp++) {
    const int p_index = prime_count - 1 - sp;
    for (int i = 0; i < p_index; i++) {
      const int64_t inv = modulus_invert((int64_t)c->q[p_index], (int64_t)c->q[i]);
      c->p_inverse_mult_r[(size_t)sp * prime_count + i] =
          modulus_multiply(inv, R_VALUE, (int64_t)c->q[i]);
    }
  }
  /* src/ckks/precomputed.cpp:123-137 */
  c->rescale_scale = calloc((size_t)prime_count * prime_count, sizeof(int64_t));
  for (int from = 0; from < prime_count; from++)
    for (int to = from + 1; to < prime_count; to++) {
      const int64_t inv = modulus_invert((int64_t)c->q[from], (int64_t)c->q[to]);
      c->rescale_scale[(size_t)from * prime_count + to] =
          modulus_multiply(inv, R_VALUE, (int64_t)c->q[to]);
    }
  return c;
}

void oracle_ctx_destroy(oracle_ctx *c) {
  if (!c)
    return;
  free(c->q); free(c->k); free(c->q2); free(c->qh);
  free(c->r2); free(c->ninv); free(c->root); free(c->root_inv);
  free(c->p_product_mult_r); free(c->p_inverse_mult_r);
  free(c->rescale_scale);
  free(c);
}

void oracle_prime_constants(const oracle_ctx *c, int i, int64_t *out,
                            int64_t *root, int64_t *root_inverse) {
  out[0] = (int64_t)c->k[i];
  out[1] = c->r2[i];
  out[2] = c->ninv[i];
  if (root)
    memcpy(root, c->root + (size_t)i * (c->n - 1), sizeof(int64_t) * (c->n - 1));
  if (root_inverse)
    memcpy(root_inverse, c->root_inv + (size_t)i * (c->n - 1),
           sizeof(int64_t) * (c->n - 1));
}

void oracle_ntt_limb(const oracle_ctx *c, int i, int64_t *x) {
  oracle_ntt_raw(x, c->root + (size_t)i * (c->n - 1), c->q[i], c->k[i], c->log_n);
}

void oracle_intt_limb(const oracle_ctx *c, int i, int64_t *x) {
  oracle_intt_raw(x, c->root_inv + (size_t)i * (c->n - 1), c->q[i], c->k[i],
                  c->log_n);
}

#define LIMB(base, idx) ((base) + (size_t)(idx) * (size_t)n)

/* src/ckks/core_ntt.cpp:91-100: mont_enter(R^2) then ntt */
void oracle_to_ntt(const oracle_ctx *c, int mid, int limbs, int64_t *x) {
  const int n = c->n;
  for (int l = 0; l < limbs; l++) {
    const int pi = mid + l;
    int64_t *y = LIMB(x, l);
    for (int j = 0; j < n; j++)
      y[j] = MONT(y[j], c->r2[pi], c->q[pi], c->k[pi]);
    oracle_ntt_limb(c, pi, y);
  }
}

/* src/ckks/core_ntt.cpp:131-150: intt then mont_enter_exit(N^-1 R)
 * (src/montgomery/compound.cpp:54-60) */
void oracle_to_intt(const oracle_ctx *c, int mid, int limbs, int64_t *x) {
  const int n = c->n;
  for (int l = 0; l < limbs; l++) {
    const int pi = mid + l;
    int64_t *y = LIMB(x, l);
    oracle_intt_limb(c, pi, y);
    for (int j = 0; j < n; j++)
      y[j] = oracle_mont_exit(MONT(y[j], c->ninv[pi], c->q[pi], c->k[pi]),
                              c->q[pi], c->k[pi]);
  }
}

/* SingleProcessor::mont_mult, src/processor/single.cpp:60-74 */
void oracle_mont_mult_limbs(const oracle_ctx *c, int mid, int limbs,
                            const int64_t *a, const int64_t *b, int64_t *out) {
  const int n = c->n;
  for (int l = 0; l < limbs; l++) {
    const int pi = mid + l;
    for (int j = 0; j < n; j++) {
      const size_t o = (size_t)l * n + j;
      out[o] = MONT(a[o], b[o], c->q[pi], c->k[pi]);
    }
  }
}

/* src/ckks/core.cpp:284-314 */
void oracle_tensor(const oracle_ctx *c, int mid, int limbs, const int64_t *a,
                   const int64_t *b, int64_t *trio) {
  const int n = c->n;
  const size_t poly = (size_t)limbs * n;
  for (int l = 0; l < limbs; l++) {
    const int pi = mid + l;
    const uint64_t q = c->q[pi], k = c->k[pi];
    const int64_t q2 = (int64_t)c->q2[pi];
    for (int j = 0; j < n; j++) {
      const size_t o = (size_t)l * n + j;
      const int64_t a0 = a[o], a1 = a[poly + o], b0 = b[o], b1 = b[poly + o];
      trio[o] = MONT(a0, b0, q, k);
      const int64_t z1 = MONT(a0, b1, q, k);
      const int64_t z3 = MONT(a1, b0, q, k);
      trio[poly + o] = guard(guard(z1 + z3, q2), (int64_t)q);
      trio[2 * poly + o] = MONT(a1, b1, q, k);
    }
  }
}

/* src/ckks/parameters.cpp:325-360 */
static int partition_count(const oracle_ctx *c, int mid) {
  const int chain = c->chain_count - mid;
  return (chain + c->alpha - 1) / c->alpha;
}

int oracle_partitions(const oracle_ctx *c, int mid, int key_mid, int *starts,
                      int *sizes, int *key_index) {
  const int chain = c->chain_count - mid;
  const int count = partition_count(c, mid);
  int pidx = partition_count(c, key_mid) - count;
  int o = 0;
  for (int start = chain - count * c->alpha; start < chain;
       start += c->alpha, pidx++, o++) {
    const int lo = start > 0 ? start : 0;
    const int hi = start + c->alpha < chain ? start + c->alpha : chain;
    starts[o] = lo;
    sizes[o] = hi - lo;
    key_index[o] = pidx;
  }
  return count;
}

/* (prod_{t<j} q[part[t]]) mod modulus, as BigInt % modulus
 * (src/ckks/precomputed.cpp:325-382) */
static int64_t partial_product_mod(const oracle_ctx *c, int mid, int start,
                                   int j, int64_t modulus) {
  int64_t r = 1 % modulus;
  for (int t = 0; t < j; t++)
    r = modulus_multiply(r, (int64_t)(c->q[mid + start + t] % (uint64_t)modulus),
                         modulus);
  return r;
}

/* Exact hybrid key switching of x (INTT form, `chain` limbs at mid)
 * following src/ckks/core_key_switch.cpp:293-358 (alpha > 1) and the
 * single-special-prime path :162-291 (alpha == 1), which computes the same
 * limbs through mont_enter_tiled + NTT per partition. */
void oracle_key_switch(const oracle_ctx *c, int mid, const int64_t *x,
                       const int64_t *key, int key_mid, int64_t *out) {
  const int n = c->n;
  const int alpha = c->alpha;
  const int chain = c->chain_count - mid;
  const int ext = chain + alpha;
  const int key_chain = c->chain_count - key_mid + alpha;
  const int key_offset = (c->chain_count - key_mid) - chain;
  int starts[128], sizes[128], kidx[128];
  const int parts = oracle_partitions(c, mid, key_mid, starts, sizes, kidx);

  int64_t *acc = calloc((size_t)2 * ext * n, sizeof(int64_t));
  int64_t *mrs = malloc(sizeof(int64_t) * (size_t)alpha * n);
  int64_t *extended = malloc(sizeof(int64_t) * (size_t)ext * n);

  for (int p = 0; p < parts; p++) {
    const int s = starts[p], size = sizes[p];
    /* rns_to_mrs, src/ckks/core_key_switch.cpp:21-56 */
    for (int j = 0; j < size; j++)
      memcpy(LIMB(mrs, j), LIMB(x, s), sizeof(int64_t) * n);
    for (int j = 0; j + 1 < size; j++) {
      const int pi = mid + s + j + 1;
      const int64_t m = (int64_t)c->q[pi];
      const int64_t qp = partial_product_mod(c, mid, s, j + 1, m);
      const int64_t inv_r = modulus_multiply(modulus_invert(qp, m), R_VALUE, m);
      const int64_t q2 = (int64_t)c->q2[pi];
      int64_t *d = LIMB(mrs, j + 1);
      const int64_t *xs = LIMB(x, s + j + 1);
      for (int t = 0; t < n; t++) {
        const int64_t diff = guard(xs[t] + q2 - d[t], q2);
        d[t] = guard(MONT(diff, inv_r, c->q[pi], c->k[pi]), m);
      }
      for (int tgt = j + 2; tgt < size; tgt++) {
        const int ti = mid + s + tgt;
        const int64_t tm = (int64_t)c->q[ti];
        const int64_t cst = modulus_multiply(
            partial_product_mod(c, mid, s, j + 1, tm), R_VALUE, tm);
        const int64_t tq2 = (int64_t)c->q2[ti];
        int64_t *y = LIMB(mrs, tgt);
        for (int t = 0; t < n; t++)
          y[t] = guard(guard(MONT(d[t], cst, c->q[ti], c->k[ti]) + y[t], tq2),
                       tm);
      }
    }
    /* mod_up, src/ckks/core_key_switch.cpp:92-108 */
    for (int i = 0; i < ext; i++) {
      const int pi = mid + i;
      const uint64_t q = c->q[pi], k = c->k[pi];
      const int64_t q2 = (int64_t)c->q2[pi];
      int64_t *y = LIMB(extended, i);
      for (int t = 0; t < n; t++)
        y[t] = MONT(mrs[t], c->r2[pi], q, k);
      for (int j = 1; j < size; j++) {
        const int64_t cst = modulus_multiply(
            partial_product_mod(c, mid, s, j, (int64_t)q), c->r2[pi], (int64_t)q);
        const int64_t *d = LIMB(mrs, j);
        for (int t = 0; t < n; t++)
          y[t] = guard(MONT(d[t], cst, q, k) + y[t], q2);
      }
      oracle_ntt_limb(c, pi, y);
    }
    /* key inner product, src/ckks/core_key_switch.cpp:318-336 */
    for (int half = 0; half < 2; half++) {
      const int64_t *kp =
          key + ((size_t)(2 * kidx[p] + half) * key_chain + key_offset) * n;
      int64_t *a = acc + (size_t)half * ext * n;
      for (int i = 0; i < ext; i++) {
        const int pi = mid + i;
        const uint64_t q = c->q[pi], k = c->k[pi];
        const int64_t q2 = (int64_t)c->q2[pi];
        const int64_t *kl = LIMB(kp, i), *e = LIMB(extended, i);
        int64_t *al = LIMB(a, i);
        if (p == 0) {
          for (int t = 0; t < n; t++)
            al[t] = MONT(e[t], kl[t], q, k);
        } else {
          for (int t = 0; t < n; t++)
            al[t] = guard(guard(MONT(e[t], kl[t], q, k) + al[t], q2),
                          (int64_t)q);
        }
      }
    }
  }

  /* mod_down, src/ckks/core_key_switch.cpp:110-160 */
  for (int half = 0; half < 2; half++) {
    int64_t *a = acc + (size_t)half * ext * n;
    oracle_to_intt(c, mid, ext, a);
    for (int i = 0; i < ext; i++) {
      const int64_t q = (int64_t)c->q[mid + i];
      int64_t *al = LIMB(a, i);
      for (int t = 0; t < n; t++)
        al[t] = guard(al[t], q);
    }
    int reduced = ext;
    for (int sp = 0; sp < alpha; sp++) {
      reduced--;
      const int64_t *xp = LIMB(a, reduced);
      for (int i = 0; i < reduced; i++) {
        const int pi = mid + i;
        const int64_t cst = c->p_inverse_mult_r[(size_t)sp * c->prime_count + pi];
        const int64_t q2 = (int64_t)c->q2[pi];
        int64_t *al = LIMB(a, i);
        for (int t = 0; t < n; t++)
          al[t] = MONT(guard(al[t] + q2 - xp[t], q2), cst, c->q[pi], c->k[pi]);
      }
    }
    int64_t *o = out + (size_t)half * chain * n;
    for (int i = 0; i < chain; i++) {
      const int64_t q = (int64_t)c->q[mid + i];
      const int64_t *al = LIMB(a, i);
      int64_t *ol = LIMB(o, i);
      for (int t = 0; t < n; t++)
        ol[t] = al[t] < 0 ? al[t] + q : guard(al[t], q);
    }
  }
  free(acc);
  free(mrs);
  free(extended);
}

/* src/ckks/core.cpp:713-741 */
void oracle_relinearize(const oracle_ctx *c, int mid, const int64_t *trio,
                        const int64_t *key, int key_mid, int64_t *duo) {
  const int n = c->n;
  const int chain = c->chain_count - mid;
  const size_t poly = (size_t)chain * n;
  int64_t *x2 = malloc(sizeof(int64_t) * poly);
  int64_t *ks = malloc(sizeof(int64_t) * 2 * poly);
  memcpy(duo, trio, sizeof(int64_t) * 2 * poly);
  oracle_to_intt(c, mid, chain, duo);
  oracle_to_intt(c, mid, chain, duo + poly);
  memcpy(x2, trio + 2 * poly, sizeof(int64_t) * poly);
  oracle_to_intt(c, mid, chain, x2);
  oracle_key_switch(c, mid, x2, key, key_mid, ks);
  for (int h = 0; h < 2; h++)
    for (int l = 0; l < chain; l++) {
      const int64_t q = (int64_t)c->q[mid + l], q2 = (int64_t)c->q2[mid + l];
      int64_t *y = duo + h * poly + (size_t)l * n;
      const int64_t *z = ks + h * poly + (size_t)l * n;
      for (int t = 0; t < n; t++)
        y[t] = guard(guard(y[t] + z[t], q2), q);
    }
  free(x2);
  free(ks);
}

/* src/ckks/core_rescale.cpp:5-25 with the processor primitives of
 * src/processor/single.cpp:76-124 */
void oracle_rescale(const oracle_ctx *c, int mid, int polys, const int64_t *x,
                    int64_t *y) {
  const int n = c->n;
  const int chain = c->chain_count - mid;
  const int next = chain - 1;
  const uint64_t qh0 = c->qh[mid];
  for (int p = 0; p < polys; p++) {
    const int64_t *xp = x + (size_t)p * chain * n;
    int64_t *yp = y + (size_t)p * next * n;
    const int64_t *x0 = xp;
    for (int i = 0; i < next; i++) {
      const int pi = mid + 1 + i;
      const int64_t q = (int64_t)c->q[pi], q2 = (int64_t)c->q2[pi];
      const int64_t cst = c->rescale_scale[(size_t)mid * c->prime_count + pi];
      const int64_t *xi = LIMB(xp, i + 1);
      int64_t *yi = LIMB(yp, i);
      for (int t = 0; t < n; t++) {
        /* get_rounder compares int64 against the unsigned q_half
         * (src/montgomery/unit.cpp:67-71): negative inputs round up */
        const int64_t rounder = (uint64_t)x0[t] > qh0 ? 1 : 0;
        int64_t v = xi[t] - x0[t];
        v = v >= 0 ? v : v + q2;
        v = guard(v, q);
        v = MONT(v, cst, c->q[pi], c->k[pi]);
        v = guard(v + rounder, q2);
        yi[t] = guard(v, q);
      }
    }
  }
}

/* src/ckks/engine_multiply.cpp:149-162 on rescaled INTT inputs, followed by
 * src/ckks/engine_rescale.cpp:50-60 */
void oracle_hmult_relin_rescale(const oracle_ctx *c, int mid, const int64_t *a,
                                const int64_t *b, const int64_t *key,
                                int key_mid, int64_t *out) {
  const int n = c->n;
  const int chain = c->chain_count - mid;
  const size_t duo = (size_t)2 * chain * n;
  int64_t *an = malloc(sizeof(int64_t) * duo);
  int64_t *bn = malloc(sizeof(int64_t) * duo);
  int64_t *trio = malloc(sizeof(int64_t) * (duo + duo / 2));
  int64_t *rel = malloc(sizeof(int64_t) * duo);
  memcpy(an, a, sizeof(int64_t) * duo);
  memcpy(bn, b, sizeof(int64_t) * duo);
  for (int p = 0; p < 2; p++) {
    oracle_to_ntt(c, mid, chain, an + (size_t)p * chain * n);
    oracle_to_ntt(c, mid, chain, bn + (size_t)p * chain * n);
  }
  oracle_tensor(c, mid, chain, an, bn, trio);
  oracle_relinearize(c, mid, trio, key, key_mid, rel);
  oracle_rescale(c, mid, 2, rel, out);
  free(an);
  free(bn);
  free(trio);
  free(rel);
}

/* src/ckks/encode.cpp:44-56: 5^(2N - shift) mod 2N, or shift itself for the
 * conjugation shift 2N - 1 */
static int64_t automorphism_jump(int n, int delta) {
  const int64_t dsize = 2 * (int64_t)n;
  const int64_t shift = ((int64_t)delta + dsize) % dsize;
  if (shift == dsize - 1)
    return shift;
  int64_t jump = 1;
  for (int64_t p = 0; p < dsize - shift; p++) {
    jump *= 5;
    jump &= dsize - 1;
  }
  return jump;
}

/* CKKSCore::rotate_ (src/ckks/core.cpp:845-869) with the permutation and
 * sign of src/ckks/encode.cpp:513-526: y[(j i mod 2N) mod N] = +-x[i], then
 * make_unsigned_reduce_2q_to_q */
void oracle_automorphism(const oracle_ctx *c, int mid, int limbs, int polys,
                         int delta, const int64_t *x, int64_t *y) {
  const int n = c->n;
  const int64_t dsize = 2 * (int64_t)n;
  const int64_t jump = automorphism_jump(n, delta);
  for (int p = 0; p < polys; p++)
    for (int l = 0; l < limbs; l++) {
      const int64_t q = (int64_t)c->q[mid + l];
      const int64_t *xl = x + ((size_t)p * limbs + l) * n;
      int64_t *yl = y + ((size_t)p * limbs + l) * n;
      for (int i = 0; i < n; i++) {
        const int64_t pi = jump * i % dsize;
        int64_t v = ((pi / n) & 1) ? -xl[i] : xl[i];
        yl[pi % n] = v < 0 ? v + q : guard(v, q);
      }
    }
}

/* key switching of x followed by ModDown and the add of d (INTT forms):
 * out0 = reduce(d + ModDown(ks0)), out1 = ModDown(ks1) */
static void switch_and_add(const oracle_ctx *c, int mid, const int64_t *x,
                           const int64_t *d, const int64_t *key, int key_mid,
                           int64_t *out) {
  const int n = c->n;
  const int chain = c->chain_count - mid;
  oracle_key_switch(c, mid, x, key, key_mid, out);
  for (int l = 0; l < chain; l++) {
    const int64_t q = (int64_t)c->q[mid + l], q2 = (int64_t)c->q2[mid + l];
    int64_t *y = out + (size_t)l * n;
    const int64_t *dl = d + (size_t)l * n;
    for (int t = 0; t < n; t++)
      y[t] = guard(guard(y[t] + dl[t], q2), q);
  }
}

/* CKKSCore::switch_key (src/ckks/core.cpp:1058-1083) on an INTT-form duo */
void oracle_switch_key(const oracle_ctx *c, int mid, const int64_t *ct,
                       const int64_t *key, int key_mid, int64_t *out) {
  const size_t poly = (size_t)(c->chain_count - mid) * c->n;
  switch_and_add(c, mid, ct + poly, ct, key, key_mid, out);
}

/* CKKSCore::rotate (src/ckks/core.cpp:954-962): decompose c1, key switcher
 * with b*P added in the NTT domain (src/ckks/core_key_switch.cpp:509-548),
 * ModDown (src/ckks/core.cpp:916-936), then the automorphism.  The b*P term
 * divided by P is exactly b, so it is added after ModDown here; the
 * canonical outputs are identical. */
void oracle_rotate(const oracle_ctx *c, int mid, const int64_t *ct,
                   const int64_t *key, int key_mid, int delta, int64_t *out) {
  const int n = c->n;
  const int chain = c->chain_count - mid;
  const size_t poly = (size_t)chain * n;
  int64_t *sw = malloc(sizeof(int64_t) * 2 * poly);
  switch_and_add(c, mid, ct + poly, ct, key, key_mid, sw);
  oracle_automorphism(c, mid, chain, 2, delta, sw, out);
  free(sw);
}

/* CKKSCore::conjugate (src/ckks/core.cpp:1035-1056): automorphism with
 * delta = -1 first, then key switching of the permuted c1 and the add of
 * the permuted c0. */
void oracle_conjugate(const oracle_ctx *c, int mid, const int64_t *ct,
                      const int64_t *key, int key_mid, int64_t *out) {
  const int n = c->n;
  const int chain = c->chain_count - mid;
  const size_t poly = (size_t)chain * n;
  int64_t *perm = malloc(sizeof(int64_t) * 2 * poly);
  oracle_automorphism(c, mid, chain, 2, -1, ct, perm);
  switch_and_add(c, mid, perm + poly, perm, key, key_mid, out);
  free(perm);
}

/* Per-coefficient primitives of one limb, indexed like the C ABI's FHE_EW_*
 * codes: src/montgomery/unit.cpp:5-86 (mont_mult, mont_enter,
 * reduce_2q_to_q, make_signed, negate, mont_add, mont_subtract,
 * get_rounder), src/montgomery/compound.cpp:6-109 (the enter / sub / exit /
 * reduce compounds, make_unsigned_reduce_2q_to_q, negate_make_unsigned) and
 * src/operation/scalar.cpp:5-11 (add_scalar).  y may alias x1 or x2; the
 * mult-add form accumulates into y. */
void oracle_elementwise_limb(int op, const int64_t *x1, const int64_t *x2,
                             int64_t *y, int n, uint64_t q, uint64_t k,
                             int64_t c) {
  const int64_t qq = (int64_t)q, q2 = (int64_t)(2 * q);
  const int64_t qh = (int64_t)(q >> 1);
  for (int i = 0; i < n; i++) {
    const int64_t a = x1[i], b = x2 ? x2[i] : 0;
    int64_t r;
    switch (op) {
    case 0: r = MONT(a, b, q, k); break;
    case 1: r = MONT(a, c, q, k); break;
    case 2: r = guard(a, qq); break;
    /* make_signed and get_rounder compare against an unsigned q/2, so a
     * negative int64 operand counts as large */
    case 3: r = (uint64_t)a <= (uint64_t)qh ? a : a - qq; break;
    case 4: r = -a; break;
    case 5: r = guard(a + b, q2); break;
    case 6: r = a - b; r = r >= 0 ? r : r + q2; break;
    case 7: r = guard(MONT(a, c, q, k) + b, q2); break;
    case 8: r = guard(guard(MONT(a, c, q, k) + b, q2), qq); break;
    case 9: r = guard(MONT(a, b, q, k) + y[i], q2); break;
    case 10: r = guard(MONT(guard(a + q2 - b, q2), c, q, k), qq); break;
    case 11: r = MONT(guard(a + q2 - b, q2), c, q, k); break;
    case 12: r = oracle_mont_exit(MONT(a, c, q, k), q, k); break;
    case 13: r = guard(oracle_mont_exit(MONT(a, c, q, k), q, k), qq); break;
    case 14: r = guard(guard(a + b, q2), qq); break;
    case 15: r = guard(guard(a + q2 - b, q2), qq); break;
    case 16: r = a < 0 ? a + qq : guard(a, qq); break;
    case 17: r = qq - a; break;
    case 18: r = i == 0 ? a + c : a; break;
    case 19: r = (uint64_t)a > (uint64_t)qh ? 1 : 0; break;
    /* mont_enter_scalar_ then reduce_2q_to_q_ (src/ckks/core.cpp:228-251);
     * add_scalar_ then reduce_2q_to_q_ (src/ckks/core.cpp:340-373) */
    case 20: r = guard(MONT(a, c, q, k), qq); break;
    case 21: r = guard(i == 0 ? a + c : a, qq); break;
    default: r = 0;
    }
    y[i] = r;
  }
}

/* make_unsigned_reduce_2q_to_q of one signed value (src/montgomery/
 * compound.cpp:85-91), as tile_unsigned_reduce_2q_to_q_ applies it to a
 * single-limb sample on every limb (src/ckks/core_calculate.cpp:485-531) */
static int64_t unsigned_reduce(int64_t x, int64_t q) {
  return x < 0 ? x + q : guard(x, q);
}

/* CKKSCore::encrypt(pt, pk) (src/ckks/core.cpp:35-82) with the host's
 * samples u (binary), e0, e1 (N signed values each) given explicitly:
 * c_h = to_intt(to_ntt(tile(u)) * pk_h) + err_h, err0 = pt + tile(e0),
 * err1 = tile(e1), every sum through mont_add_reduce_2q_to_q.  pk is the
 * public key's [2][pk_limbs][N] from pk_moduli_id; out is [2][limbs][N]. */
void oracle_encrypt_pk(const oracle_ctx *c, int mid, const int64_t *pt,
                       const int64_t *pk, int pk_mid, const int64_t *u,
                       const int64_t *e0, const int64_t *e1, int64_t *out) {
  const int n = c->n;
  const int limbs = c->chain_count - mid;
  const int pk_limbs = c->chain_count - pk_mid;
  const size_t poly = (size_t)limbs * n;
  int64_t *un = malloc(sizeof(int64_t) * poly);
  for (int l = 0; l < limbs; l++)
    for (int j = 0; j < n; j++)
      un[(size_t)l * n + j] = unsigned_reduce(u[j], (int64_t)c->q[mid + l]);
  oracle_to_ntt(c, mid, limbs, un);
  for (int h = 0; h < 2; h++) {
    const int64_t *pkh = pk + ((size_t)h * pk_limbs + (mid - pk_mid)) * n;
    oracle_mont_mult_limbs(c, mid, limbs, un, pkh, out + h * poly);
    oracle_to_intt(c, mid, limbs, out + h * poly);
  }
  for (int l = 0; l < limbs; l++) {
    const int64_t q = (int64_t)c->q[mid + l], q2 = (int64_t)c->q2[mid + l];
    for (int j = 0; j < n; j++) {
      const size_t o = (size_t)l * n + j;
      const int64_t err0 =
          guard(guard(pt[o] + unsigned_reduce(e0[j], q), q2), q);
      const int64_t err1 = unsigned_reduce(e1[j], q);
      out[o] = guard(guard(out[o] + err0, q2), q);
      out[poly + o] = guard(guard(out[poly + o] + err1, q2), q);
    }
  }
  free(un);
}

/* CKKSCore::decrypt_duo (src/ckks/core.cpp:171-194): pt = c0 +
 * to_intt(to_ntt(c1) * s), the sum through mont_add_reduce_2q_to_q.  sk is
 * the secret key's [sk_limbs][N] from sk_moduli_id. */
void oracle_decrypt(const oracle_ctx *c, int mid, const int64_t *ct,
                    const int64_t *sk, int sk_mid, int64_t *out) {
  const int n = c->n;
  const int limbs = c->chain_count - mid;
  const size_t poly = (size_t)limbs * n;
  memcpy(out, ct + poly, sizeof(int64_t) * poly);
  oracle_to_ntt(c, mid, limbs, out);
  oracle_mont_mult_limbs(c, mid, limbs, out, sk + (size_t)(mid - sk_mid) * n,
                         out);
  oracle_to_intt(c, mid, limbs, out);
  for (int l = 0; l < limbs; l++) {
    const int64_t q = (int64_t)c->q[mid + l], q2 = (int64_t)c->q2[mid + l];
    for (int j = 0; j < n; j++) {
      const size_t o = (size_t)l * n + j;
      out[o] = guard(guard(ct[o] + out[o], q2), q);
    }
  }
}

/* ---- encoding ------------------------------------------------------------
 * The double arithmetic below is written one IEEE operation at a time in the
 * reference's order; this file is built as ISO C for x86-64, where nothing is
 * contracted into fused multiply-adds, as in the reference's build. */

static const double kPi = 3.14159265358979323846;

/* glibc's sin and sincos through volatile pointers, unfoldable: the
 * reference's -O2 build computes every same-argument (cos x, sin x) pair
 * with sincos (its results differ from separate calls in the last bit for
 * some x) and sin(-x) beside such a pair with a separate sin call
 * (src/ckks/precomputed.cpp:67-75, src/ckks/encode.cpp:264-297). */
static double (*volatile libm_sin)(double) = sin;
static void (*volatile libm_sincos)(double, double *, double *) = sincos;

/* Lists the cycles of a permutation, each from its smallest unvisited start
 * (src/ckks/encode.cpp:89-124). */
static void permutation_cycles(const uint64_t *perm, int size, uint64_t *out) {
  char *seen = calloc((size_t)size, 1);
  int w = 0, start = 0;
  while (start < size) {
    uint64_t cur = perm[start];
    uint64_t nxt = perm[cur];
    for (;;) {
      out[w++] = cur;
      seen[cur] = 1;
      cur = nxt;
      if (seen[nxt])
        break;
      nxt = perm[cur];
    }
    start++;
    while (start < size && seen[start])
      start++;
  }
  free(seen);
}

/* src/ckks/encode.cpp:24-150: post = conjugate(fold, shift), pre = post^-1,
 * with shift the circular shift of the two halves and fold the map
 * i -> (g (2 i + 1) mod 2 codec - 1) / 2, g = 5^(2 codec - 1) mod 2 codec. */
void oracle_encode_pre_permutation(int codec, uint64_t *pre) {
  if (codec < 2)
    return;
  const int half = codec / 2;
  const int64_t two = 2 * (int64_t)codec;
  uint64_t *shift = malloc(sizeof(uint64_t) * codec);
  uint64_t *fold = malloc(sizeof(uint64_t) * codec);
  uint64_t *c1 = malloc(sizeof(uint64_t) * codec);
  uint64_t *c2 = malloc(sizeof(uint64_t) * codec);
  uint64_t *post = malloc(sizeof(uint64_t) * codec);
  shift[0] = (uint64_t)(half - 1);
  for (int i = 1; i < half; i++)
    shift[i] = (uint64_t)(i - 1);
  for (int i = half; i < codec - 1; i++)
    shift[i] = (uint64_t)(i + 1);
  shift[codec - 1] = (uint64_t)half;
  int64_t g = 1;
  for (int64_t e = 0; e < two - 1; e++)
    g = (g * 5) % two;
  for (int i = 0; i < codec; i++)
    fold[i] = (uint64_t)(((g * (2 * (int64_t)i + 1)) % two - 1) / 2);
  permutation_cycles(fold, codec, c1);
  permutation_cycles(shift, codec, c2);
  for (int i = 0; i < codec; i++)
    post[c1[i]] = c2[i];
  for (int i = 0; i < codec; i++)
    pre[post[i]] = (uint64_t)i;
  free(shift);
  free(fold);
  free(c1);
  free(c2);
  free(post);
}

/* positive_modulo (src/arithmetic/arithmetic.cpp:271-273) */
static int64_t positive_mod(int64_t x, int64_t q) { return (q + x % q) % q; }

/* CKKSCore::encode_polynomial (src/ckks/core_encode.cpp:31-150) with
 * encode_to_double_ (:11-29): pre_permute (src/ckks/encode.cpp:156-185),
 * bit_reversal_permute and fft (src/operation/fft.cpp:57-102), normalize
 * (:50-55), apply_twister (src/ckks/encode.cpp:264-275), then llround +
 * tile_unsigned_reduce (src/processor/single.cpp:555-566, :315-326) or
 * double_to_rns above 2^59 (src/arithmetic/arithmetic.cpp:242-269), and
 * mont_enter with r_mult (src/montgomery/unit.cpp:12-17). */
void oracle_encode(const oracle_ctx *c, const double *values,
                   int complex_input, double scale_before_tile, int first,
                   int enc_mid, int out_limbs, const int64_t *r_mult,
                   int64_t *out) {
  const int codec = c->n, slots = c->n / 2;
  uint64_t *pre = malloc(sizeof(uint64_t) * codec);
  double *re = malloc(sizeof(double) * codec);
  double *im = malloc(sizeof(double) * codec);
  double *v = malloc(sizeof(double) * codec);
  oracle_encode_pre_permutation(codec, pre);
  for (int i = 0; i < slots; i++) {
    const int64_t p = (int64_t)pre[i];
    const double a = complex_input ? values[2 * i] : values[i];
    const double b = complex_input ? values[2 * i + 1] : 0.0;
    re[p] = a;
    im[p] = b;
    re[codec - 1 - p] = a;
    im[codec - 1 - p] = complex_input ? -b : 0.0;
  }
  /* bit reversal: index i <-> its log2(codec)-bit reverse */
  int log_codec = 0;
  while ((1 << log_codec) < codec)
    log_codec++;
  for (int i = 0; i < codec; i++) {
    int r = 0;
    for (int b = 0; b < log_codec; b++)
      if (i & (1 << b))
        r |= 1 << (log_codec - 1 - b);
    if (i < r) {
      double t = re[i];
      re[i] = re[r];
      re[r] = t;
      t = im[i];
      im[i] = im[r];
      im[r] = t;
    }
  }
  /* iterative radix-2 decimation in time with w_t = exp(-2 pi i t / codec) */
  const double unit_angle = -2 * kPi / codec;
  for (int step = 2, tstep = codec / 2; step <= codec; step <<= 1, tstep >>= 1) {
    const int h = step >> 1;
    for (int off = 0; off < codec; off += step)
      for (int j = 0; j < h; j++) {
        const double angle = (j * tstep) * unit_angle;
        double wr, wi;
        libm_sincos(angle, &wi, &wr);
        const int x = off + j, y = off + j + h;
        const double pr = wr * re[y] - wi * im[y];
        const double pi_ = wr * im[y] + wi * re[y];
        const double lr = re[x], li = im[x];
        re[x] = lr + pr;
        im[x] = li + pi_;
        re[y] = lr - pr;
        im[y] = li - pi_;
      }
  }
  /* normalize, twist, scale */
  const double twist_step = -kPi / (codec * 2);
  double max_abs = 0.0;
  for (int i = 0; i < codec; i++) {
    const double a = re[i] / codec, b = im[i] / codec;
    const double ang = (2 * i) * twist_step;
    double sn, cs;
    libm_sincos(ang, &sn, &cs);
    v[i] = (a * cs - b * sn) * scale_before_tile;
    if (fabs(v[i]) > max_abs)
      max_abs = fabs(v[i]);
  }
  const int big = max_abs > (double)(UINT64_C(1) << 59);
  for (int l = 0; l < out_limbs; l++) {
    int64_t *o = out + (size_t)l * codec;
    if (l < first) {
      memset(o, 0, sizeof(int64_t) * codec);
      continue;
    }
    const int e = l - first, p = enc_mid + e;
    const int64_t q = (int64_t)c->q[p];
    for (int j = 0; j < codec; j++) {
      int64_t y;
      if (big) {
        if (fabs(v[j]) < (double)(INT64_C(1) << 62)) {
          y = positive_mod((int64_t)(v[j] + 0.5), q);
        } else {
          int exponent;
          const double m = frexp(v[j] + 0.5, &exponent);
          y = positive_mod((int64_t)(m * (double)(INT64_C(1) << 52)), q);
          for (int le = 52; le < exponent; le++) {
            y <<= 1;
            y %= q;
          }
        }
      } else {
        y = unsigned_reduce(llround(v[j]), q);
      }
      o[j] = oracle_mont_mult(y, r_mult[e], c->q[p], c->k[p]);
    }
  }
  free(pre);
  free(re);
  free(im);
  free(v);
}

/* ---- decoding ------------------------------------------------------------
 * Multiword integers are little-endian uint64 arrays with an explicit length
 * (BigInt, src/arithmetic/big_int.cpp, keeps its words trimmed of high
 * zeros; comparisons here are numeric, which is what its trimmed >= is). */

typedef struct {
  int len;
  uint64_t w[80];
} mw;

static void mw_trim(mw *a) {
  while (a->len > 1 && a->w[a->len - 1] == 0)
    a->len--;
}

static void mw_set(mw *a, uint64_t v) {
  memset(a, 0, sizeof(*a));
  a->len = 1;
  a->w[0] = v;
}

static void mw_mul_word(const mw *a, uint64_t b, mw *out) {
  mw r;
  memset(&r, 0, sizeof(r));
  u128 carry = 0;
  for (int i = 0; i < a->len; i++) {
    const u128 p = (u128)a->w[i] * b + carry;
    r.w[i] = (uint64_t)p;
    carry = p >> 64;
  }
  r.w[a->len] = (uint64_t)carry;
  r.len = a->len + 1;
  mw_trim(&r);
  *out = r;
}

static void mw_add(const mw *a, const mw *b, mw *out) {
  mw r;
  memset(&r, 0, sizeof(r));
  const int len = a->len > b->len ? a->len : b->len;
  u128 carry = 0;
  for (int i = 0; i < len; i++) {
    const u128 s = (u128)(i < a->len ? a->w[i] : 0) +
                   (i < b->len ? b->w[i] : 0) + carry;
    r.w[i] = (uint64_t)s;
    carry = s >> 64;
  }
  r.w[len] = (uint64_t)carry;
  r.len = len + 1;
  mw_trim(&r);
  *out = r;
}

static int mw_geq(const mw *a, const mw *b) {
  if (a->len != b->len)
    return a->len > b->len;
  for (int i = a->len - 1; i >= 0; i--)
    if (a->w[i] != b->w[i])
      return a->w[i] > b->w[i];
  return 1;
}

/* a - b for a >= b */
static void mw_sub(const mw *a, const mw *b, mw *out) {
  mw r;
  memset(&r, 0, sizeof(r));
  uint64_t borrow = 0;
  for (int i = 0; i < a->len; i++) {
    const uint64_t bw = i < b->len ? b->w[i] : 0;
    const uint64_t d = a->w[i] - bw;
    r.w[i] = d - borrow;
    borrow = (a->w[i] < bw) | (d < borrow);
  }
  r.len = a->len;
  mw_trim(&r);
  *out = r;
}

static uint64_t mw_div_word(const mw *a, uint64_t b, mw *out) {
  mw r;
  memset(&r, 0, sizeof(r));
  u128 rem = 0;
  for (int i = a->len - 1; i >= 0; i--) {
    const u128 cur = (rem << 64) | a->w[i];
    r.w[i] = (uint64_t)(cur / b);
    rem = cur % b;
  }
  r.len = a->len;
  mw_trim(&r);
  if (out)
    *out = r;
  return (uint64_t)rem;
}

static int64_t inverse_mod(int64_t a, int64_t m) {
  /* a^(m-2) mod m for prime m */
  int64_t r = 1, b = a % m, e = m - 2;
  while (e > 0) {
    if (e & 1)
      r = (int64_t)((i128)r * b % m);
    b = (int64_t)((i128)b * b % m);
    e >>= 1;
  }
  return r;
}

/* the CRT lift of one plaintext at moduli_id to doubles
 * (src/ckks/core_decode.cpp:68-98, src/arithmetic/arithmetic.cpp:169-197) */
void oracle_decode_crt(const oracle_ctx *c, const int64_t *cur, int mid,
                       double *v) {
  const int n = c->n;
  const int limbs = c->chain_count - mid;
  if (limbs == 1) {
    const int64_t q = (int64_t)c->q[mid];
    for (int j = 0; j < n; j++) {
      const int64_t x = cur[j];
      v[j] = (double)((uint64_t)x <= c->qh[mid] ? x : x - q);
    }
  } else {
    /* DecodeParameters: Q, Q/2, reversers, -Q^-1 mod 2^62 */
    mw Q, half, upper, *rev = malloc(sizeof(mw) * limbs);
    mw_set(&Q, 1);
    for (int i = 0; i < limbs; i++)
      mw_mul_word(&Q, c->q[mid + i], &Q);
    mw_div_word(&Q, 2, &half);
    for (int i = 0; i < limbs; i++) {
      const uint64_t qi = c->q[mid + i];
      mw cof;
      mw_div_word(&Q, qi, &cof);
      uint64_t inv = 1;
      for (int j = 0; j < limbs; j++)
        if (j != i)
          inv = (uint64_t)((u128)inv *
                           (uint64_t)inverse_mod((int64_t)(c->q[mid + j] % qi),
                                                 (int64_t)qi) % qi);
      mw_mul_word(&cof, inv, &rev[i]);
    }
    const uint64_t r62 = UINT64_C(1) << 62, mask = r62 - 1;
    uint64_t qinv = Q.w[0];
    for (int it = 0; it < 6; it++)
      qinv *= 2 - Q.w[0] * qinv;
    const uint64_t kq = (0 - qinv) & mask;
    mw_mul_word(&Q, r62, &upper);
    for (int j = 0; j < n; j++) {
      mw acc;
      mw_set(&acc, 0);
      for (int i = 0; i < limbs; i++) {
        const int64_t q = (int64_t)c->q[mid + i];
        int64_t x = cur[(size_t)i * n + j];
        x = x < 0 ? x + q : guard(x, q);
        const uint64_t xr = (uint64_t)((u128)(uint64_t)x * r62 % (uint64_t)q);
        mw t;
        mw_mul_word(&rev[i], xr, &t);
        mw_add(&acc, &t, &acc);
      }
      const uint64_t w0 = acc.w[0];
      while (mw_geq(&acc, &upper))
        mw_sub(&acc, &upper, &acc);
      mw prod, quo;
      mw_mul_word(&Q, (w0 * kq) & mask, &prod);
      mw_add(&acc, &prod, &acc);
      /* exact division by 2^62 */
      memset(&quo, 0, sizeof(quo));
      for (int w = 0; w < acc.len; w++)
        quo.w[w] = (acc.w[w] >> 62) | (w + 1 < acc.len ? acc.w[w + 1] << 2 : 0);
      quo.len = acc.len;
      mw_trim(&quo);
      if (mw_geq(&quo, &Q))
        mw_sub(&quo, &Q, &quo);
      int negative = 0;
      if (mw_geq(&quo, &half)) {
        mw_sub(&Q, &quo, &quo);
        negative = 1;
      }
      double y = 0;
      for (int w = 0; w < quo.len; w++)
        y += (double)quo.w[w] * pow(2.0, 64 * w);
      v[j] = negative ? -y : y;
    }
    free(rev);
  }
}

/* src/ckks/core_decode.cpp:10-124 with rns_to_double
 * (src/arithmetic/arithmetic.cpp:169-197), DecodeParameters
 * (src/parameter/decode.cpp:5-51), BigInt::montgomery_modulus and get_double
 * (src/arithmetic/big_int.cpp:202-225), apply_skewer
 * (src/ckks/encode.cpp:289-297), the FFT with the inverse twiddles and
 * permute_after_ifft (src/ckks/encode.cpp:481-503). */
int oracle_decode(const oracle_ctx *c, const int64_t *pt, int mid,
                  long double ckks_scale, double base_scale,
                  int complex_output, double *out) {
  const int n = c->n, codec = c->n, slots = c->n / 2;
  const int full_level = c->chain_count - 2;
  int limbs = c->chain_count - mid;
  if (limbs == 1)
    return -1;
  long double modulus = 1;
  for (int i = 0; i < limbs; i++)
    modulus *= (long double)c->q[mid + i];
  if (ckks_scale > modulus)
    return -1;
  int64_t *cur = malloc(sizeof(int64_t) * (size_t)limbs * n);
  memcpy(cur, pt, sizeof(int64_t) * (size_t)limbs * n);
  long double scale = ckks_scale;
  while (mid <= full_level) {
    const long double rescaled = scale / c->q[mid];
    if (rescaled <= base_scale)
      break;
    int64_t *next = malloc(sizeof(int64_t) * (size_t)(limbs - 1) * n);
    oracle_rescale(c, mid, 1, cur, next);
    free(cur);
    cur = next;
    scale = rescaled;
    mid++;
    limbs--;
  }
  double *v = malloc(sizeof(double) * codec);
  oracle_decode_crt(c, cur, mid, v);
  free(cur);
  /* skewer, bit reversal, FFT with (cos a, sin -a), gather */
  double *re = malloc(sizeof(double) * codec);
  double *im = malloc(sizeof(double) * codec);
  const double skew_step = kPi / (codec * 2);
  for (int i = 0; i < codec; i++) {
    const double a = (2 * i) * skew_step;
    double sn, cs;
    libm_sincos(a, &sn, &cs);
    re[i] = v[i] * cs;
    im[i] = v[i] * sn;
  }
  int log_codec = 0;
  while ((1 << log_codec) < codec)
    log_codec++;
  for (int i = 0; i < codec; i++) {
    int r = 0;
    for (int b = 0; b < log_codec; b++)
      if (i & (1 << b))
        r |= 1 << (log_codec - 1 - b);
    if (i < r) {
      double t = re[i];
      re[i] = re[r];
      re[r] = t;
      t = im[i];
      im[i] = im[r];
      im[r] = t;
    }
  }
  const double unit_angle = -2 * kPi / codec;
  for (int step = 2, tstep = codec / 2; step <= codec; step <<= 1, tstep >>= 1) {
    const int h = step >> 1;
    for (int off = 0; off < codec; off += step)
      for (int j = 0; j < h; j++) {
        const double angle = (j * tstep) * unit_angle;
        double wr, unused;
        libm_sincos(angle, &unused, &wr);
        const double wi = libm_sin(-angle);
        const int x = off + j, y = off + j + h;
        const double pr = wr * re[y] - wi * im[y];
        const double pi_ = wr * im[y] + wi * re[y];
        const double lr = re[x], li = im[x];
        re[x] = lr + pr;
        im[x] = li + pi_;
        re[y] = lr - pr;
        im[y] = li - pi_;
      }
  }
  uint64_t *pre = malloc(sizeof(uint64_t) * codec);
  oracle_encode_pre_permutation(codec, pre);
  const double scale_inverse = (double)(1.0 / scale);
  for (int i = 0; i < slots; i++) {
    /* post[index] = i  <=>  index = pre[i] */
    const uint64_t index = pre[i];
    if (complex_output) {
      out[2 * i] = re[index] * scale_inverse;
      out[2 * i + 1] = im[index] * scale_inverse;
    } else {
      out[i] = re[index] * scale_inverse;
    }
  }
  free(pre);
  free(re);
  free(im);
  free(v);
  return 0;
}
