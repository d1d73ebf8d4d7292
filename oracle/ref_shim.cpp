// Test-only C wrapper around the reference CKKS engine.
//
// This file is oracle infrastructure: it is compiled together with the
// reference's own sources (by oracle/Makefile, outputs under oracle/_ref/) and
// is loaded only by tests/, the smoke check and bench.py's cpu_baseline leg.
// It never ships in the product library.
//
// Every entry point forwards to the reference's public API
// (src/engine/abstract_engine.h, src/ckks/engine.h) or to its free functions
// (src/operation/ntt.h, src/ckks/precomputed.h).  Ciphertexts and keys are
// exchanged as plain int64 arrays in the reference's limb order
// [polynomial][limb][coefficient].
#include <complex>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "arithmetic/arithmetic.h"
#include "ckks/config.h"
#include "ckks/core.h"
#include "ckks/encode.h"
#include "ckks/engine.h"
#include "ckks/precomputed.h"
#include "engine/engine.h"
#include "montgomery/compound.h"
#include "montgomery/unit.h"
#include "operation/scalar.h"
#include "parameter/multiplier.h"
#include "operation/ntt.h"
#include "operation/fft.h"
#include "operation/bootstrap.h"
#include "arithmetic/pi.h"

using namespace desilo::fhe;

namespace {

thread_local std::string g_error;

struct RefEngine {
  CKKSConfig config;
  CKKSEngine *engine;
  CKKSPrecomputed *precomputed;
};

CKKSConfig make_config(int log_n, int special_primes, int bootstrap_preset,
                       int thread_count) {
  CKKSConfig config;
  config.log_coeff_count_ = log_n;
  config.special_prime_count_ = special_primes;
  config.scheme_preset_ = bootstrap_preset ? SchemePreset::CKKS_54_60_BOOTSTRAP
                                           : SchemePreset::CKKS_40_60_DEFAULT;
  if (thread_count > 1) {
    config.processor_ = EngineProcessor::CPU_PARALLEL;
    config.thread_count_ = thread_count;
  }
  return config;
}

void copy_out(Data *data, int64_t *out) {
  const int polys = data->polynomial_count();
  const int chains = data->chain_count_;
  const int n = data->coeff_count_;
  for (int p = 0; p < polys; p++) {
    auto poly = data->get_polynomial(p);
    for (int c = 0; c < chains; c++) {
      std::memcpy(out + ((size_t)p * chains + c) * n, poly->get(c),
                  sizeof(int64_t) * n);
    }
  }
}

void copy_in(Data *data, const int64_t *in) {
  const int polys = data->polynomial_count();
  const int chains = data->chain_count_;
  const int n = data->coeff_count_;
  for (int p = 0; p < polys; p++) {
    auto poly = data->get_polynomial(p);
    for (int c = 0; c < chains; c++) {
      std::memcpy(poly->get(c), in + ((size_t)p * chains + c) * n,
                  sizeof(int64_t) * n);
    }
  }
}

template <class F> auto guarded(F &&f) -> decltype(f()) {
  try {
    return f();
  } catch (const std::exception &e) {
    g_error = e.what();
  } catch (...) {
    g_error = "unknown exception";
  }
  return decltype(f())();
}

} // namespace

extern "C" {

const char *ref_last_error() { return g_error.c_str(); }

void *ref_engine_create(int log_n, int special_primes, int bootstrap_preset,
                        int thread_count) {
  return guarded([&]() -> void * {
    auto config =
        make_config(log_n, special_primes, bootstrap_preset, thread_count);
    config.validate();
    auto ref = new RefEngine{config, new CKKSEngine(config),
                             new CKKSPrecomputed(config)};
    return ref;
  });
}

void ref_engine_destroy(void *handle) {
  auto ref = static_cast<RefEngine *>(handle);
  if (!ref)
    return;
  delete ref->engine;
  delete ref->precomputed;
  delete ref;
}

// out[0..8] = N, prime_count, chain_count, special_prime_count, full_level,
//             max_level, slot_count, scale_bits, one_more_level
int ref_engine_info(void *handle, int64_t *out) {
  return guarded([&]() -> int {
    auto ref = static_cast<RefEngine *>(handle);
    const auto &p = ref->precomputed->parameters_;
    out[0] = p.get_coeff_count();
    out[1] = p.get_prime_count();
    out[2] = p.get_chain_count();
    out[3] = p.get_special_prime_count();
    out[4] = p.get_full_level();
    out[5] = p.get_max_level();
    out[6] = p.get_slot_count();
    out[7] = p.get_scale_bit_length();
    out[8] = p.get_use_one_more_level();
    return 0;
  });
}

int ref_engine_primes(void *handle, uint64_t *out) {
  return guarded([&]() -> int {
    auto ref = static_cast<RefEngine *>(handle);
    const auto prime_count = ref->precomputed->prime_count_;
    const auto q = ref->precomputed->get_q(0);
    for (int i = 0; i < prime_count; i++)
      out[i] = q[i];
    return prime_count;
  });
}

int ref_moduli_id(void *handle, int level) {
  return guarded([&]() -> int {
    auto ref = static_cast<RefEngine *>(handle);
    return ref->precomputed->parameters_.get_moduli_id(level);
  });
}

// Per-prime derived constants, for checking a restatement of
// src/ckks/derived_parameters.cpp:19-50 and src/ckks/precomputed.cpp:33-217.
// out = {k, r_squared, coeff_count_inverse}, root/root_inverse get N-1 each.
int ref_prime_constants(void *handle, int prime_index, int64_t *out,
                        int64_t *root, int64_t *root_inverse) {
  return guarded([&]() -> int {
    auto ref = static_cast<RefEngine *>(handle);
    auto pre = ref->precomputed;
    const int n = pre->coeff_count_;
    out[0] = (int64_t)pre->get_k(0)[prime_index];
    out[1] = pre->get_r_squared(0)[prime_index];
    out[2] = pre->get_coeff_count_inverse(0)[prime_index];
    if (root)
      std::memcpy(root, pre->get_root(0, prime_index),
                  sizeof(int64_t) * (n - 1));
    if (root_inverse)
      std::memcpy(root_inverse, pre->get_root_inverse(0, prime_index),
                  sizeof(int64_t) * (n - 1));
    return 0;
  });
}

// In-place forward / inverse NTT of one limb with the reference's free
// functions (src/operation/ntt.cpp:46-134, :177-272).
int ref_ntt_limb(void *handle, int prime_index, int64_t *x, int inverse) {
  return guarded([&]() -> int {
    auto ref = static_cast<RefEngine *>(handle);
    auto pre = ref->precomputed;
    const int n = pre->coeff_count_;
    const uint64_t q = pre->get_q(0)[prime_index];
    const uint64_t k = pre->get_k(0)[prime_index];
    const uint64_t q2 = pre->get_q_double(0)[prime_index];
    if (inverse)
      intt(x, x, pre->get_root_inverse(0, prime_index), q2, q, k, n,
           pre->log_coeff_count_);
    else
      ntt(x, x, pre->get_root(0, prime_index), q2, q, k, n,
          pre->log_coeff_count_);
    return 0;
  });
}

void *ref_secret_key(void *handle) {
  return guarded([&]() -> void * {
    return static_cast<RefEngine *>(handle)->engine->create_secret_key();
  });
}

void *ref_public_key(void *handle, void *sk) {
  return guarded([&]() -> void * {
    return static_cast<RefEngine *>(handle)->engine->create_public_key(
        static_cast<SecretKey *>(sk));
  });
}

void *ref_relin_key(void *handle, void *sk) {
  return guarded([&]() -> void * {
    return static_cast<RefEngine *>(handle)->engine->create_relinearization_key(
        static_cast<SecretKey *>(sk));
  });
}

void ref_free_secret_key(void *p) { delete static_cast<SecretKey *>(p); }
void ref_free_public_key(void *p) { delete static_cast<PublicKey *>(p); }
void ref_free_relin_key(void *p) { delete static_cast<RelinearizationKey *>(p); }
void ref_free_ct(void *p) { delete static_cast<UnitCiphertext *>(p); }

// shape[0..3] = moduli_id, polynomial_count, chain_count, coeff_count
static void data_shape(Data *d, int64_t *shape) {
  shape[0] = d->moduli_id_;
  shape[1] = d->polynomial_count();
  shape[2] = d->chain_count_;
  shape[3] = d->coeff_count_;
}

int ref_relin_key_shape(void *rlk, int64_t *shape) {
  data_shape(static_cast<RelinearizationKey *>(rlk)->key_switching_key_, shape);
  return 0;
}
int ref_relin_key_get(void *rlk, int64_t *out) {
  copy_out(static_cast<RelinearizationKey *>(rlk)->key_switching_key_, out);
  return 0;
}
int ref_relin_key_set(void *rlk, const int64_t *in) {
  copy_in(static_cast<RelinearizationKey *>(rlk)->key_switching_key_, in);
  return 0;
}

// shape[0..5] = moduli_id, polynomial_count, chain_count, coeff_count, flag,
//               level
int ref_ct_shape(void *ct, int64_t *shape) {
  auto c = static_cast<UnitCiphertext *>(ct);
  data_shape(c, shape);
  shape[4] = static_cast<int64_t>(c->flag_);
  shape[5] = c->level_;
  return 0;
}
double ref_ct_scale(void *ct) {
  return (double)static_cast<UnitCiphertext *>(ct)->ckks_scale_;
}
int ref_ct_get(void *ct, int64_t *out) {
  copy_out(static_cast<UnitCiphertext *>(ct), out);
  return 0;
}
int ref_ct_set(void *ct, const int64_t *in) {
  copy_in(static_cast<UnitCiphertext *>(ct), in);
  return 0;
}

// Public-key encryption of real slot values at a level (level < 0: max).
void *ref_encrypt(void *handle, const double *values, int count, void *pk,
                  int level) {
  return guarded([&]() -> void * {
    auto ref = static_cast<RefEngine *>(handle);
    std::vector<double> input(values, values + count);
    auto key = static_cast<PublicKey *>(pk);
    if (level < 0)
      return ref->engine->encrypt_unit(input, key);
    return ref->engine->encrypt_unit(input, key, level);
  });
}

int ref_decrypt(void *handle, void *ct, void *sk, double *out) {
  return guarded([&]() -> int {
    auto ref = static_cast<RefEngine *>(handle);
    auto values = ref->engine->decrypt(static_cast<UnitCiphertext *>(ct),
                                       static_cast<SecretKey *>(sk));
    std::memcpy(out, values.data(), sizeof(double) * values.size());
    return (int)values.size();
  });
}

void *ref_multiply_relinearize(void *handle, void *a, void *b, void *rlk) {
  return guarded([&]() -> void * {
    auto ref = static_cast<RefEngine *>(handle);
    return ref->engine->multiply(static_cast<UnitCiphertext *>(a),
                                 static_cast<UnitCiphertext *>(b),
                                 static_cast<RelinearizationKey *>(rlk));
  });
}

void *ref_rescale(void *handle, void *ct) {
  return guarded([&]() -> void * {
    auto ref = static_cast<RefEngine *>(handle);
    return ref->engine->rescale(static_cast<UnitCiphertext *>(ct));
  });
}

// Rotation / conjugation keys (src/ckks/engine.cpp:82-124) and their
// KeySwitchingKey limbs.
void *ref_fixed_rotation_key(void *handle, void *sk, int delta) {
  return guarded([&]() -> void * {
    return static_cast<RefEngine *>(handle)->engine->create_fixed_rotation_key(
        static_cast<SecretKey *>(sk), delta);
  });
}
void *ref_conjugation_key(void *handle, void *sk) {
  return guarded([&]() -> void * {
    return static_cast<RefEngine *>(handle)->engine->create_conjugation_key(
        static_cast<SecretKey *>(sk));
  });
}
void ref_free_fixed_rotation_key(void *p) {
  delete static_cast<FixedRotationKey *>(p);
}
void ref_free_conjugation_key(void *p) {
  delete static_cast<ConjugationKey *>(p);
}
int ref_fixed_rotation_key_delta(void *frk) {
  return static_cast<FixedRotationKey *>(frk)->delta_;
}
static Data *ksk_of(void *key, int kind) {
  if (kind == 0)
    return static_cast<RelinearizationKey *>(key)->key_switching_key_;
  if (kind == 1)
    return static_cast<FixedRotationKey *>(key)->key_switching_key_;
  if (kind == 2)
    return static_cast<ConjugationKey *>(key)->key_switching_key_;
  if (kind == 3)
    return static_cast<SecretKey *>(key);
  if (kind == 5)
    return static_cast<SmallBootstrapKey *>(key)->to_sparse_;
  if (kind == 6)
    return static_cast<SmallBootstrapKey *>(key)->to_dense_;
  return static_cast<PublicKey *>(key);
}
// kind: 0 relinearization, 1 fixed rotation, 2 conjugation, 3 secret key,
// 4 public key
int ref_key_shape(void *key, int kind, int64_t *shape) {
  data_shape(ksk_of(key, kind), shape);
  return 0;
}
int ref_key_get(void *key, int kind, int64_t *out) {
  copy_out(ksk_of(key, kind), out);
  return 0;
}
int ref_key_set(void *key, int kind, const int64_t *in) {
  copy_in(ksk_of(key, kind), in);
  return 0;
}

// CKKSEngine::rotate(ct, FixedRotationKey) (src/ckks/engine_rotate.cpp:78-101)
void *ref_rotate(void *handle, void *ct, void *frk) {
  return guarded([&]() -> void * {
    auto ref = static_cast<RefEngine *>(handle);
    return ref->engine->rotate(static_cast<UnitCiphertext *>(ct),
                               static_cast<FixedRotationKey *>(frk));
  });
}

// CKKSEngine::conjugate(ct, ConjugationKey) (src/ckks/engine_conjugate.cpp:53-73)
void *ref_conjugate(void *handle, void *ct, void *cjk) {
  return guarded([&]() -> void * {
    auto ref = static_cast<RefEngine *>(handle);
    return ref->engine->conjugate(static_cast<UnitCiphertext *>(ct),
                                  static_cast<ConjugationKey *>(cjk));
  });
}

// Automorphism tables of src/ckks/encode.cpp:513-557 for N coefficients.
int ref_permutation_and_sign(int n, int delta, int32_t *permutation,
                             int32_t *sign) {
  return guarded([&]() -> int {
    get_permutation_and_sign(permutation, sign, n, delta);
    return 0;
  });
}

int ref_ntt_permutation(int n, int delta, int32_t *permutation) {
  return guarded([&]() -> int {
    get_ntt_permutation(permutation, n, delta);
    return 0;
  });
}

// One limb of a per-coefficient primitive through the reference's own limb
// functions (src/montgomery/unit.cpp, src/montgomery/compound.cpp,
// src/operation/scalar.cpp), with the op codes of the C ABI's FHE_EW_*.
// The accumulate forms start from y = x2 (enter-add) or the given y.
int ref_elementwise(int op, const int64_t *x1_in, const int64_t *x2_in,
                    int64_t *y, int n, uint64_t q, uint64_t k, int64_t c) {
  return guarded([&]() -> int {
    std::vector<int64_t> x1(x1_in, x1_in + n), x2(n, 0);
    if (x2_in)
      std::copy(x2_in, x2_in + n, x2.begin());
    const uint64_t q2 = q << 1, qh = q >> 1;
    int64_t *a = x1.data(), *b = x2.data();
    switch (op) {
    case 0: mont_mult(a, b, y, q, k, n); break;
    case 1: mont_enter(a, y, c, q, k, n); break;
    case 2: reduce_2q_to_q(a, y, q, n); break;
    case 3: make_signed(a, y, q, qh, n); break;
    case 4: negate(a, y, 1, n); break;
    case 5: mont_add(a, b, y, q2, n); break;
    case 6: mont_subtract(a, b, y, q2, n); break;
    case 7: std::copy(b, b + n, y); mont_enter_add(a, y, c, q2, q, k, n); break;
    case 8:
      std::copy(b, b + n, y);
      mont_enter_add_reduce_2q_to_q(a, y, c, q2, q, k, n);
      break;
    case 9: mont_mult_add(a, b, y, q2, q, k, n); break;
    case 10: mont_sub_enter_reduce_2q_to_q(a, b, y, c, q2, q, k, n); break;
    case 11: mont_sub_enter(a, b, y, c, q2, q, k, n); break;
    case 12: mont_enter_exit(a, y, c, q, k, n); break;
    case 13: mont_enter_exit_reduce_2q_to_q(a, y, c, q, k, n); break;
    case 14: mont_add_reduce_2q_to_q(a, b, y, q2, q, n); break;
    case 15: mont_sub_reduce_2q_to_q(a, b, y, q2, q, n); break;
    case 16: make_unsigned_reduce_2q_to_q(a, y, q, n); break;
    case 17: negate_make_unsigned(a, y, q, n); break;
    case 18: add_scalar(a, y, c, n); break;
    case 19: get_rounder(a, y, qh, n); break;
    case 20:
      mont_enter(a, y, c, q, k, n);
      reduce_2q_to_q(y, y, q, n);
      break;
    case 21:
      add_scalar(a, y, c, n);
      reduce_2q_to_q(y, y, q, n);
      break;
    default: throw std::invalid_argument("unknown op");
    }
    return 0;
  });
}

// CKKSPrecomputed's multiplier scalars (src/ckks/precomputed.cpp:418-533).
// kind 0 add(value, level), 1 mult(value, level), 2 mult(int_value,
// moduli_id = level), 3 level_down(level, level2), 4
// level_down_after_rescale(level, level2).  Returns the scalar count.
int ref_multiplier(void *handle, int kind, double value, int64_t int_value,
                   int level, int level2, int64_t *out) {
  return guarded([&]() -> int {
    auto pre = static_cast<RefEngine *>(handle)->precomputed;
    Multiplier *m = nullptr;
    switch (kind) {
    case 0: m = pre->make_add_multiplier(value, level); break;
    case 1: m = pre->make_mult_multiplier(value, level); break;
    case 2: m = pre->make_mult_multiplier(int_value, level); break;
    case 3: m = pre->make_level_down_multiplier(level, level2); break;
    case 4:
      m = pre->make_level_down_multiplier_after_rescale(level, level2);
      break;
    default: throw std::invalid_argument("unknown multiplier kind");
    }
    const auto values = m->get();
    const int count = int(m->multiplier_.size());
    for (int i = 0; i < count; i++)
      out[i] = values[i];
    delete m;
    return count;
  });
}

// Scalar operations and level down through the reference engine
// (src/ckks/engine_multiply.cpp:418-457, src/ckks/engine_add.cpp:518-536,
// src/ckks/engine_level_down.cpp:5-37, src/ckks/engine_negate.cpp).
void *ref_multiply_double(void *handle, void *ct, double value) {
  return guarded([&]() -> void * {
    return static_cast<RefEngine *>(handle)->engine->multiply(
        static_cast<UnitCiphertext *>(ct), value);
  });
}

void *ref_multiply_int(void *handle, void *ct, int64_t value) {
  return guarded([&]() -> void * {
    return static_cast<RefEngine *>(handle)->engine->multiply(
        static_cast<UnitCiphertext *>(ct), value);
  });
}

void *ref_add_double(void *handle, void *ct, double value) {
  return guarded([&]() -> void * {
    return static_cast<RefEngine *>(handle)->engine->add(
        static_cast<UnitCiphertext *>(ct), value);
  });
}

void *ref_level_down(void *handle, void *ct, int level) {
  return guarded([&]() -> void * {
    return static_cast<RefEngine *>(handle)->engine->level_down(
        static_cast<UnitCiphertext *>(ct), level);
  });
}

void *ref_add_ct(void *handle, void *a, void *b) {
  return guarded([&]() -> void * {
    return static_cast<RefEngine *>(handle)->engine->add(
        static_cast<UnitCiphertext *>(a), static_cast<UnitCiphertext *>(b));
  });
}

void *ref_subtract_ct(void *handle, void *a, void *b) {
  return guarded([&]() -> void * {
    return static_cast<RefEngine *>(handle)->engine->subtract(
        static_cast<UnitCiphertext *>(a), static_cast<UnitCiphertext *>(b));
  });
}

void *ref_negate(void *handle, void *ct) {
  return guarded([&]() -> void * {
    return static_cast<RefEngine *>(handle)->engine->negate(
        static_cast<UnitCiphertext *>(ct));
  });
}

// CKKSEngine::decrypt_to_unit_plaintext (src/ckks/engine_decrypt.cpp):
// the plaintext limbs before decoding.  shape as ref_ct_shape.
int ref_decrypt_limbs(void *handle, void *ct, void *sk, int64_t *shape,
                      int64_t *out) {
  return guarded([&]() -> int {
    auto ref = static_cast<RefEngine *>(handle);
    auto pt = ref->engine->decrypt_to_unit_plaintext(
        static_cast<UnitCiphertext *>(ct), static_cast<SecretKey *>(sk));
    data_shape(pt, shape);
    if (out)
      copy_out(pt, out);
    delete pt;
    return 0;
  });
}

// CKKSEngine::encode_unit (src/ckks/engine_encode.cpp:10-36): values are
// slot_count doubles, or slot_count {re, im} pairs when complex_input.
// shape as ref_ct_shape; out [chain][N] when given.
int ref_encode(void *handle, const double *values, int complex_input,
               int level, int64_t *shape, int64_t *out) {
  return guarded([&]() -> int {
    auto ref = static_cast<RefEngine *>(handle);
    const int slots = ref->precomputed->slot_count_;
    UnitPlaintext *pt = nullptr;
    if (complex_input) {
      std::vector<std::complex<double>> v(slots);
      for (int i = 0; i < slots; i++)
        v[i] = {values[2 * i], values[2 * i + 1]};
      pt = ref->engine->encode_unit(v, level);
    } else {
      std::vector<double> v(values, values + slots);
      pt = ref->engine->encode_unit(v, level);
    }
    data_shape(pt, shape);
    if (out)
      copy_out(pt, out);
    delete pt;
    return 0;
  });
}

// CKKSPrecomputed's per-level encoding constants
// (src/ckks/precomputed.cpp:138-215); r_mult gets the level's
// r_mult_scale_after_tile, *count entries: the chain with the special primes,
// one less at the top of a one-more-level chain (:161-170).
int ref_encode_constants(void *handle, int level, long double *ckks_scale,
                         double *scale_before_tile,
                         int64_t *log_scale_after_tile, int64_t *r_mult,
                         int *count) {
  return guarded([&]() -> int {
    auto ref = static_cast<RefEngine *>(handle);
    const auto *pc = ref->precomputed;
    const auto &params = pc->parameters_;
    const int mid = params.get_moduli_id(level);
    const long double cs = pc->get_ckks_scale(level);
    const auto q = pc->get_q(mid);
    const bool top = pc->one_more_level_ && mid == 0 && (cs >= q[0] * pc->scale_);
    const int n = int(params.get_chain_count(mid, true)) - (top ? 1 : 0);
    const int64_t *r = pc->get_r_mult_scale_after_tile(level);
    for (int i = 0; i < n; i++)
      r_mult[i] = r[i];
    *count = n;
    *ckks_scale = cs;
    *scale_before_tile = pc->get_scale_before_tile(level);
    *log_scale_after_tile = pc->get_log_scale_after_tile(level);
    return top ? 1 : 0;
  });
}

// CKKSEngine::decode_unit / decode_complex_unit_plaintext
// (src/ckks/engine_decode.cpp:6-18) of a plaintext at `level` with the given
// limbs [chain][N] and long double ckks_scale.  out: slot_count doubles or
// {re, im} pairs.
int ref_decode(void *handle, const int64_t *limbs, int level,
               const long double *ckks_scale, int complex_output,
               double *out) {
  return guarded([&]() -> int {
    auto ref = static_cast<RefEngine *>(handle);
    const int slots = ref->precomputed->slot_count_;
    std::vector<double> zeros(slots, 0.0);
    auto shape = ref->engine->encode_unit(zeros, level);
    auto src = shape->get_polynomial(0);
    auto poly = new Polynomial(src->flag_, src->moduli_id_, shape->chain_count_,
                               shape->coeff_count_);
    std::vector<Polynomial *> polys{poly};
    auto pt = new UnitPlaintext(
        polys, shape->flag_, shape->scheme_preset_, shape->log_slot_count_,
        shape->log_coeff_count_, shape->special_prime_count_,
        shape->moduli_id_, shape->level_, shape->chain_count_,
        shape->coeff_count_, *ckks_scale);
    delete shape;
    copy_in(pt, limbs);
    if (complex_output) {
      auto v = ref->engine->decode_complex_unit_plaintext(pt);
      for (int i = 0; i < slots; i++) {
        out[2 * i] = v[i].real();
        out[2 * i + 1] = v[i].imag();
      }
    } else {
      auto v = ref->engine->decode_unit(pt);
      std::memcpy(out, v.data(), sizeof(double) * slots);
    }
    delete pt;
    return slots;
  });
}

// The CRT step of the decoder alone: rns_to_double
// (src/arithmetic/arithmetic.cpp:169-197) of limbs [chain][N] at moduli_id
// after make_unsigned_reduce (src/ckks/core_decode.cpp:77-98).
int ref_rns_to_double(void *handle, const int64_t *limbs, int moduli_id,
                      double *out) {
  return guarded([&]() -> int {
    auto ref = static_cast<RefEngine *>(handle);
    const auto *pc = ref->precomputed;
    const int chain = pc->parameters_.get_chain_count(moduli_id);
    const int n = pc->coeff_count_;
    const auto q = pc->get_q(moduli_id);
    std::vector<std::vector<int64_t>> x(chain, std::vector<int64_t>(n));
    std::vector<int64_t *> rows(chain);
    for (int i = 0; i < chain; i++) {
      for (int j = 0; j < n; j++) {
        const int64_t v = limbs[(size_t)i * n + j];
        x[i][j] = v < 0 ? v + int64_t(q[i]) : guard(v, int64_t(q[i]));
      }
      rows[i] = x[i].data();
    }
    auto dp = pc->make_decode_parameters(chain, moduli_id);
    rns_to_double(rows.data(), out, q, dp->q_product_, dp->q_product_half_,
                  dp->rns_reverser_.data(), dp->r_minus_k_product_, chain, n,
                  0, n);
    delete dp;
    return 0;
  });
}

// The decoder after its CRT step (CKKSCore::decode_,
// src/ckks/core_decode.cpp:101-112): skewer, bit reversal, FFT with the
// inverse twiddles, permute_after_ifft with scale_inverse.  v: N doubles.
int ref_decode_tail(void *handle, const double *v, double scale_inverse,
                    int complex_output, double *out) {
  return guarded([&]() -> int {
    auto ref = static_cast<RefEngine *>(handle);
    const auto *pc = ref->precomputed;
    const int codec = pc->codec_coeff_count_, slots = pc->slot_count_;
    std::vector<double> x(v, v + codec), y(2 * codec);
    apply_skewer(x.data(), y.data(), codec, 0, pi / (codec * 2));
    auto c = reinterpret_cast<std::complex<double> *>(y.data());
    bit_reversal_permute(c, c, codec);
    fft(c, c, pc->ifft_twiddle_factor_.data(), codec);
    if (complex_output)
      permute_after_ifft(y.data(), reinterpret_cast<std::complex<double> *>(out),
                         pc->get_post_permutation(), slots, scale_inverse, 0,
                         codec);
    else
      permute_after_ifft(y.data(), out, pc->get_post_permutation(), slots,
                         scale_inverse, 0, codec);
    return 0;
  });
}

// CKKSCore::rescale of a plaintext (src/ckks/core_rescale.cpp:55-81): limbs
// [chain][N] at `level` -> out [chain - 1][N].
int ref_rescale_plaintext(void *handle, const int64_t *limbs, int level,
                          int64_t *out) {
  return guarded([&]() -> int {
    auto ref = static_cast<RefEngine *>(handle);
    CKKSCore core(ref->config);
    std::vector<double> zeros(ref->precomputed->slot_count_, 0.0);
    auto pt = core.encode(zeros.data(), level, false);
    copy_in(pt, limbs);
    auto r = core.rescale(pt);
    copy_out(r, out);
    delete pt;
    delete r;
    return 0;
  });
}

// CKKSEngine::create_small_bootstrap_key (src/ckks/engine.cpp:140-160):
// key-switching keys to the sparse secret (at level 0) and back (at the
// maximum key-switching level); kinds 5 and 6 of ref_key_*.
void *ref_small_bootstrap_key(void *handle, void *sk) {
  return guarded([&]() -> void * {
    auto ref = static_cast<RefEngine *>(handle);
    return ref->engine->create_small_bootstrap_key(static_cast<SecretKey *>(sk));
  });
}
void ref_free_small_bootstrap_key(void *p) {
  delete static_cast<SmallBootstrapKey *>(p);
}

// CKKSEngine::mod_raise_ (src/ckks/engine_mod_raise.cpp:5-34) of a
// ciphertext already at level 0.
void *ref_mod_raise(void *handle, void *ct, void *sbk) {
  return guarded([&]() -> void * {
    auto ref = static_cast<RefEngine *>(handle);
    return ref->engine->mod_raise_(static_cast<UnitCiphertext *>(ct),
                                   static_cast<SmallBootstrapKey *>(sbk));
  });
}

// The steps of CKKSEngine::mod_raise_ one at a time through CKKSCore
// (step 0: rescale, 1: switch_key to the sparse secret, 2: make_signed_tile,
// 3: switch_key back to the dense secret); returns the step's output.
void *ref_mod_raise_step(void *handle, void *ct, void *sbk, int step) {
  return guarded([&]() -> void * {
    auto ref = static_cast<RefEngine *>(handle);
    CKKSCore core(ref->config);
    auto c = static_cast<UnitCiphertext *>(ct);
    auto k = static_cast<SmallBootstrapKey *>(sbk);
    if (step == 0)
      return core.rescale(c);
    if (step == 1)
      return core.switch_key(c, k->to_sparse_);
    if (step == 2)
      return core.make_signed_tile(c);
    return core.switch_key(c, k->to_dense_);
  });
}

// CKKSEngine::multiply_imaginary (src/ckks/engine_multiply.cpp:468-491).
void *ref_multiply_imaginary(void *handle, void *ct, int64_t value) {
  return guarded([&]() -> void * {
    auto ref = static_cast<RefEngine *>(handle);
    return ref->engine->multiply_imaginary(static_cast<UnitCiphertext *>(ct),
                                           value);
  });
}

// The flag of a ciphertext (DataFlag bits).
int ref_ct_flag(void *ct) {
  return int(static_cast<UnitCiphertext *>(ct)->flag_);
}

// The long double ckks_scale of a ciphertext.
int ref_ct_scale_ld(void *ct, long double *out) {
  *out = static_cast<UnitCiphertext *>(ct)->ckks_scale_;
  return 0;
}

// CKKSPrecomputed's FFT twiddle tables (src/ckks/precomputed.cpp:67-75) as
// {re, im} pairs, codec entries each.
int ref_fft_twiddles(void *handle, double *fft, double *ifft) {
  return guarded([&]() -> int {
    auto ref = static_cast<RefEngine *>(handle);
    const auto *pc = ref->precomputed;
    const int codec = pc->codec_coeff_count_;
    std::memcpy(fft, pc->fft_twiddle_factor_.data(), sizeof(double) * 2 * codec);
    std::memcpy(ifft, pc->ifft_twiddle_factor_.data(),
                sizeof(double) * 2 * codec);
    return codec;
  });
}

// CKKSPrecomputed::pre_permutation_ (src/ckks/precomputed.cpp:50-58)
int ref_pre_permutation(void *handle, uint64_t *out) {
  return guarded([&]() -> int {
    auto ref = static_cast<RefEngine *>(handle);
    const auto &pre = ref->precomputed->pre_permutation_;
    std::memcpy(out, pre.data(), pre.size() * sizeof(uint64_t));
    return int(pre.size());
  });
}

// One HMult+relin+rescale unit of work through the reference engine:
// multiply (src/ckks/engine_multiply.cpp:149-162) then rescale
// (src/ckks/engine_rescale.cpp:50-60).
void *ref_hmult_relin_rescale(void *handle, void *a, void *b, void *rlk) {
  return guarded([&]() -> void * {
    auto ref = static_cast<RefEngine *>(handle);
    auto product = ref->engine->multiply(static_cast<UnitCiphertext *>(a),
                                         static_cast<UnitCiphertext *>(b),
                                         static_cast<RelinearizationKey *>(rlk));
    auto rescaled = ref->engine->rescale(product);
    delete product;
    return rescaled;
  });
}

// UnitPlaintext handles: encode_unit of real values at a level
// (src/ckks/engine_encode.cpp:10-36), their shape (as ref_ct_shape) and limbs.
void *ref_encode_pt(void *handle, const double *values, int level) {
  return guarded([&]() -> void * {
    auto ref = static_cast<RefEngine *>(handle);
    std::vector<double> v(values, values + ref->precomputed->slot_count_);
    return ref->engine->encode_unit(v, level);
  });
}
void ref_free_pt(void *p) { delete static_cast<UnitPlaintext *>(p); }
int ref_pt_shape(void *pt, int64_t *shape) {
  auto p = static_cast<UnitPlaintext *>(pt);
  data_shape(p, shape);
  shape[4] = static_cast<int64_t>(p->flag_);
  shape[5] = p->level_;
  return 0;
}
int ref_pt_get(void *pt, int64_t *out) {
  copy_out(static_cast<UnitPlaintext *>(pt), out);
  return 0;
}

// Forms of texts: CKKSEngine::ntt / rescale of a ciphertext (plain = 0) or
// a plaintext (plain = 1) (src/ckks/engine_ntt.cpp:26-42,
// src/ckks/engine_rescale.cpp:38-60).
void *ref_ntt_text(void *handle, void *p, int plain) {
  return guarded([&]() -> void * {
    auto e = static_cast<RefEngine *>(handle)->engine;
    if (plain)
      return e->ntt(static_cast<UnitPlaintext *>(p));
    return e->ntt(static_cast<UnitCiphertext *>(p));
  });
}
void *ref_rescale_text(void *handle, void *p, int plain) {
  return guarded([&]() -> void * {
    auto e = static_cast<RefEngine *>(handle)->engine;
    if (plain)
      return e->rescale(static_cast<UnitPlaintext *>(p));
    return e->rescale(static_cast<UnitCiphertext *>(p));
  });
}
void *ref_level_down_pt(void *handle, void *p, int level) {
  return guarded([&]() -> void * {
    return static_cast<RefEngine *>(handle)->engine->level_down(
        static_cast<UnitPlaintext *>(p), level);
  });
}

// CKKSEngine::multiply(ct, ct), square(ct) with or without a relinearization
// key, multiply(ct, pt) (src/ckks/engine_multiply.cpp:39-57, 86-99, 175-186,
// 210-227).
void *ref_multiply_ct(void *handle, void *a, void *b) {
  return guarded([&]() -> void * {
    return static_cast<RefEngine *>(handle)->engine->multiply(
        static_cast<UnitCiphertext *>(a), static_cast<UnitCiphertext *>(b));
  });
}
void *ref_square(void *handle, void *a, void *rlk) {
  return guarded([&]() -> void * {
    auto e = static_cast<RefEngine *>(handle)->engine;
    auto ct = static_cast<UnitCiphertext *>(a);
    if (rlk)
      return e->square(ct, static_cast<RelinearizationKey *>(rlk));
    return e->square(ct);
  });
}
void *ref_multiply_pt(void *handle, void *ct, void *pt) {
  return guarded([&]() -> void * {
    return static_cast<RefEngine *>(handle)->engine->multiply(
        static_cast<UnitCiphertext *>(ct), static_cast<UnitPlaintext *>(pt));
  });
}

// CKKSEngine::relinearize (src/ckks/engine_relinearize.cpp:13-28).
void *ref_relinearize(void *handle, void *ct, void *rlk) {
  return guarded([&]() -> void * {
    return static_cast<RefEngine *>(handle)->engine->relinearize(
        static_cast<UnitCiphertext *>(ct),
        static_cast<RelinearizationKey *>(rlk));
  });
}

// RotationKey (src/ckks/engine.cpp:125-138): its fixed rotation keys, in
// the key's own order, through the FixedRotationKey accessors above.
void *ref_rotation_key(void *handle, void *sk) {
  return guarded([&]() -> void * {
    return static_cast<RefEngine *>(handle)->engine->create_rotation_key(
        static_cast<SecretKey *>(sk));
  });
}
void ref_free_rotation_key(void *p) { delete static_cast<RotationKey *>(p); }
int ref_rotation_key_count(void *rk) {
  return int(static_cast<RotationKey *>(rk)->fixed_rotation_keys_.size());
}
void *ref_rotation_key_fixed(void *rk, int index) {
  return static_cast<RotationKey *>(rk)->fixed_rotation_keys_.at(index);
}

// CKKSEngine::evaluate_polynomial (src/ckks/engine_advanced.cpp:365-387)
// with double coefficients.
void *ref_evaluate_polynomial(void *handle, void *ct, const double *coeffs,
                              int count, void *rlk) {
  return guarded([&]() -> void * {
    std::vector<double> c(coeffs, coeffs + count);
    return static_cast<RefEngine *>(handle)->engine->evaluate_polynomial(
        static_cast<UnitCiphertext *>(ct), c,
        static_cast<RelinearizationKey *>(rlk));
  });
}

// CKKSEngine::sum_1d_by_row (src/ckks/engine_advanced.cpp:549-577) of a
// Ciphertext holding a copy of one unit ciphertext; returns a copy of the
// result's unit ciphertext.
void *ref_sum_1d_by_row(void *handle, void *ct, void *rk) {
  return guarded([&]() -> void * {
    auto e = static_cast<RefEngine *>(handle)->engine;
    auto in = new Ciphertext(std::vector<UnitCiphertext *>{
        e->clone(static_cast<UnitCiphertext *>(ct))});
    Ciphertext *out = e->sum_1d_by_row(in, static_cast<RotationKey *>(rk));
    UnitCiphertext *r = e->clone(out->get(0));
    if (out != in)
      delete out;
    delete in;
    return r;
  });
}

// CKKSEngine::rotate(ct, RotationKey, delta) (src/ckks/engine_rotate.cpp:
// 103-130): the fixed keys whose bit the normalized delta has, in order.
void *ref_rotate_rotation_key(void *handle, void *ct, void *rk, int delta) {
  return guarded([&]() -> void * {
    return static_cast<RefEngine *>(handle)->engine->rotate(
        static_cast<UnitCiphertext *>(ct), static_cast<RotationKey *>(rk),
        delta);
  });
}

// CKKSEngine::coeff_to_slot / slot_to_coeff / bootstrap with a RotationKey
// (src/ckks/engine_bootstrap.cpp:330-454, 712-754).
int ref_coeff_to_slot(void *handle, void *ct, void *conj, void *rk,
                      int stage_count, void **real, void **imag) {
  return guarded([&]() -> int {
    auto [r, i] = static_cast<RefEngine *>(handle)->engine->coeff_to_slot(
        static_cast<UnitCiphertext *>(ct), static_cast<ConjugationKey *>(conj),
        static_cast<RotationKey *>(rk), stage_count);
    *real = r;
    *imag = i;
    return 0;
  });
}
void *ref_slot_to_coeff(void *handle, void *real, void *imag, void *conj,
                        void *rk, int stage_count) {
  return guarded([&]() -> void * {
    return static_cast<RefEngine *>(handle)->engine->slot_to_coeff(
        static_cast<UnitCiphertext *>(real),
        static_cast<UnitCiphertext *>(imag),
        static_cast<ConjugationKey *>(conj), static_cast<RotationKey *>(rk),
        stage_count);
  });
}
void *ref_bootstrap(void *handle, void *ct, void *rlk, void *conj, void *rk,
                    void *sbk, int stage_count) {
  return guarded([&]() -> void * {
    return static_cast<RefEngine *>(handle)->engine->bootstrap(
        static_cast<UnitCiphertext *>(ct),
        static_cast<RelinearizationKey *>(rlk),
        static_cast<ConjugationKey *>(conj), static_cast<RotationKey *>(rk),
        static_cast<SmallBootstrapKey *>(sbk), stage_count);
  });
}

// CKKSEngine::create_bootstrap_key (src/ckks/engine.cpp:162-222) and the
// parts of the BootstrapKey (src/data/bootstrap.h:24-44,
// src/data/plain_matrix.h:10-24) a device engine uploads.
void *ref_bootstrap_key(void *handle, void *sk, int stage_count,
                        const char *size) {
  return guarded([&]() -> void * {
    return static_cast<RefEngine *>(handle)->engine->create_bootstrap_key(
        static_cast<SecretKey *>(sk), stage_count, size);
  });
}
void ref_free_bootstrap_key(void *p) { delete static_cast<BootstrapKey *>(p); }
// out[0..3] = stage_count, inner_loop_shift, rotation contexts, CtS stages
int ref_bootstrap_key_info(void *bk, int *out) {
  auto k = static_cast<BootstrapKey *>(bk);
  out[0] = k->stage_count_;
  out[1] = k->inner_loop_shift_;
  out[2] = int(k->rotation_contexts_.size());
  out[3] = int(k->coeff_to_slot_stages_.size());
  return 0;
}
// Rotation context `index` in the map's (delta) order: its delta and fixed
// rotation key (kind 1 of ref_key_*); nullptr past the end.
void *ref_bootstrap_key_rotation(void *bk, int index, int *delta) {
  auto k = static_cast<BootstrapKey *>(bk);
  int i = 0;
  for (auto &[d, ctx] : k->rotation_contexts_)
    if (i++ == index) {
      *delta = d;
      return ctx.first;
    }
  return nullptr;
}
void *ref_bootstrap_key_small(void *bk) {
  return static_cast<BootstrapKey *>(bk)->small_bootstrap_key_;
}
// One PlainMatrix: meta[0..2] = level, inner count, outer count; the inner
// and outer deltas (when the arrays are given).
int ref_bootstrap_key_stage(void *bk, int to_slot, int stage, int *meta,
                            int *inner_deltas, int *outer_deltas) {
  return guarded([&]() -> int {
    auto k = static_cast<BootstrapKey *>(bk);
    auto &stages = to_slot ? k->coeff_to_slot_stages_ : k->slot_to_coeff_stages_;
    PlainMatrix *m = stages.at(size_t(stage));
    meta[0] = m->level_;
    meta[1] = m->inner_loop_count_;
    meta[2] = m->outer_loop_count_;
    if (inner_deltas)
      for (size_t i = 0; i < m->inner_loop_deltas_.size(); i++)
        inner_deltas[i] = m->inner_loop_deltas_[i];
    if (outer_deltas)
      for (size_t i = 0; i < m->outer_loop_deltas_.size(); i++)
        outer_deltas[i] = m->outer_loop_deltas_[i];
    return 0;
  });
}
// The encoded diagonal (outer, inner) of a stage (ref_pt_shape /
// ref_pt_get), nullptr when the matrix has none there.
void *ref_bootstrap_key_diagonal(void *bk, int to_slot, int stage, int outer,
                                 int inner) {
  auto k = static_cast<BootstrapKey *>(bk);
  auto &stages = to_slot ? k->coeff_to_slot_stages_ : k->slot_to_coeff_stages_;
  if (stage < 0 || size_t(stage) >= stages.size())
    return nullptr;
  auto &diag = stages[size_t(stage)]->encoded_diagonals_;
  auto o = diag.find(outer);
  if (o == diag.end())
    return nullptr;
  auto i = o->second.find(inner);
  return i == o->second.end() ? nullptr : i->second;
}
// CKKSEngine::coeff_to_slot / slot_to_coeff / bootstrap with a BootstrapKey
// (src/ckks/engine_bootstrap.cpp:368-398, 436-454, 794-834).
int ref_coeff_to_slot_bk(void *handle, void *ct, void *conj, void *bk,
                         void **real, void **imag) {
  return guarded([&]() -> int {
    auto [r, i] = static_cast<RefEngine *>(handle)->engine->coeff_to_slot(
        static_cast<UnitCiphertext *>(ct), static_cast<ConjugationKey *>(conj),
        static_cast<BootstrapKey *>(bk));
    *real = r;
    *imag = i;
    return 0;
  });
}
void *ref_slot_to_coeff_bk(void *handle, void *real, void *imag, void *conj,
                           void *bk) {
  return guarded([&]() -> void * {
    return static_cast<RefEngine *>(handle)->engine->slot_to_coeff(
        static_cast<UnitCiphertext *>(real),
        static_cast<UnitCiphertext *>(imag),
        static_cast<ConjugationKey *>(conj), static_cast<BootstrapKey *>(bk));
  });
}
void *ref_bootstrap_bk(void *handle, void *ct, void *rlk, void *conj,
                       void *bk) {
  return guarded([&]() -> void * {
    return static_cast<RefEngine *>(handle)->engine->bootstrap(
        static_cast<UnitCiphertext *>(ct),
        static_cast<RelinearizationKey *>(rlk),
        static_cast<ConjugationKey *>(conj), static_cast<BootstrapKey *>(bk));
  });
}

// get_coeff_to_slot_twiddles / get_slot_to_coeff_twiddles
// (src/operation/bootstrap.cpp:125-263), one stage as {re, im} pairs.
int ref_bootstrap_twiddles(int log_slots, int stage_count, int to_slot,
                           int stage, double *out) {
  return guarded([&]() -> int {
    const int slots = 1 << log_slots;
    const auto tw =
        to_slot ? get_coeff_to_slot_twiddles(slots, log_slots, stage_count)
                : get_slot_to_coeff_twiddles(slots, log_slots, stage_count);
    const auto &st = tw.at(stage);
    if (out)
      for (size_t r = 0; r < st.size(); r++)
        std::memcpy(out + r * slots * 2, st[r].data(), size_t(slots) * 16);
    return int(st.size());
  });
}

// CKKSEngine::round (src/ckks/engine_bootstrap.cpp:672-691).
void *ref_round(void *handle, void *ct, void *rlk) {
  return guarded([&]() -> void * {
    return static_cast<RefEngine *>(handle)->engine->round(
        static_cast<UnitCiphertext *>(ct),
        static_cast<RelinearizationKey *>(rlk));
  });
}

} // extern "C"
