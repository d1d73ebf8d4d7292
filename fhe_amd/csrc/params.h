// Host-side CKKS parameter set: prime chain, per-prime Montgomery constants,
// NTT twiddles, level bookkeeping and key-switching partitions.
//
// The chain layout and every constant follow the reference's definitions so
// that device results match its limbs bit for bit:
//   prime chain       src/ckks/parameters.cpp:9-105, src/ckks/config.cpp:49-84,
//                     src/arithmetic/prime.cpp:840-1015
//   per-prime         src/ckks/derived_parameters.cpp:19-50
//   chain constants   src/ckks/precomputed.cpp:33-137
//   partitions        src/ckks/parameters.cpp:325-360,
//                     src/ckks/precomputed.cpp:325-382
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace fhe {

// Number theory (host, exact with 128-bit intermediates).
int64_t mod_mul(int64_t a, int64_t b, int64_t m);
int64_t mod_pow(int64_t base, int64_t e, int64_t m);
int64_t mod_inverse(int64_t x, int64_t m);  // 0 when not invertible
bool is_prime_u64(uint64_t n);
int64_t find_ntt_root(int64_t q, int n);    // primitive 2n-th root
uint64_t mont_k(uint64_t q);                // -q^-1 mod 2^62

// Prime searches that reproduce the reference's chains.
std::vector<uint64_t> decreasing_ntt_primes(int bits, int n, int count);
std::vector<uint64_t> alternating_ntt_primes(int bits, int n, int count);

struct CkksConfig {
  int log_n = 14;
  int special_prime_count = 1;
  int scale_bits = 40;       // 40 (default preset) or 54 (bootstrap preset)
  int small_prime_bits = 40; // 40 or 54
  int large_prime_bits = 60;
  bool one_more_level = true;
  int max_level = -1;
};

// One key-switching partition (digit) of a ciphertext at some moduli_id.
struct Partition {
  int start;      // first limb, relative to the ciphertext's moduli_id
  int size;       // limbs in the digit (<= special prime count)
  int key_index;  // partition index inside the key
};

class CkksParams {
public:
  // twiddles = false skips the NTT tables (host-only queries)
  static CkksParams from_config(const CkksConfig &config,
                                bool twiddles = true);
  static CkksParams from_primes(int log_n, const std::vector<uint64_t> &primes,
                                int special_prime_count, int scale_bits,
                                bool one_more_level, bool twiddles = true);

  int log_n = 0, n = 0;
  int prime_count = 0;   // Q chain + specials
  int chain_count = 0;   // Q chain only
  int alpha = 0;         // special prime count
  int full_level = 0;
  int max_level = 0;
  int scale_bits = 0;
  bool one_more_level = false;

  std::vector<uint64_t> q, k, q2, qh;
  std::vector<int64_t> r2, ninv;
  // Montgomery-form twiddles, N-1 used entries per prime, padded to N:
  // root[p*N + i] = psi^bitrev(i+1) * R, root_inverse in backward-paint order.
  std::vector<int64_t> root, root_inverse;
  // The same twiddles in standard form with their Shoup quotients,
  // interleaved per entry: [prime][N][2] = {w, floor(w * 2^64 / q)}.  They
  // drive the lazy transforms of stages whose outputs are canonicalized
  // later, where any representative of the residue is acceptable.
  std::vector<uint64_t> root_shoup, root_inverse_shoup;
  // Scale 2^scale_bits and the per-level scale deviations, index = level
  // 0..full_level (src/ckks/parameters.cpp:54-104).
  double scale = 0;
  std::vector<long double> deviation_ld;
  std::vector<double> deviation;
  std::vector<int64_t> p_product_mult_r;   // [chain_count]
  std::vector<int64_t> p_inverse_mult_r;   // [alpha][prime_count]
  // the same inverses in standard form with Shoup quotients,
  // [alpha][prime_count][2] = {p^-1 mod q, floor(p^-1 * 2^64 / q)}
  std::vector<uint64_t> p_inverse_shoup;
  // largest multiple of q below 2^63, per prime: adding it makes any
  // difference x - y with x >= 0 and |y| < 2^62 non-negative and below 2^64
  std::vector<uint64_t> q_top;
  std::vector<int64_t> rescale_scale;      // [prime_count][prime_count]

  int limbs_at(int moduli_id) const { return chain_count - moduli_id; }
  int moduli_id_of_level(int level) const { return full_level - level; }
  int level_of(int moduli_id, bool rescaled) const {
    return rescaled ? full_level - moduli_id + 1 : full_level - moduli_id;
  }
  int max_key_switching_level() const;
  // CKKSCore::get_mod_raise_ratio (src/ckks/core.cpp:1112-1119), full slots
  int64_t mod_raise_ratio() const;
  int partition_count(int moduli_id) const {
    return (limbs_at(moduli_id) + alpha - 1) / alpha;
  }
  std::vector<Partition> partitions(int moduli_id, int key_moduli_id) const;

  // Device constants for ModUp of a ciphertext at moduli_id: for each
  // partition p, a block of ks_stride int64 entries
  //   [0, 16)            Garner inverses (prod q_<j)^-1 * R mod q_j, j>=1
  //   [16, 16+256)       Garner tile constants (prod q_<=j) * R mod q_t
  //   [272, 272+16*ext)  basis-extension constants (prod q_<j) * R^2 mod q_i
  //   [272+16*ext, +32*ext)  the same in standard form (prod q_<j) * R mod
  //                      q_i with its Shoup quotient, pairs at 2 (j ext + i)
  static constexpr int kMaxAlpha = 16;
  static constexpr int kKsInvOffset = 0;
  static constexpr int kKsTileOffset = 16;
  static constexpr int kKsExtOffset = 16 + 16 * 16;
  int ks_ext_shoup_offset(int moduli_id) const {
    return kKsExtOffset + kMaxAlpha * (limbs_at(moduli_id) + alpha);
  }
  int ks_stride(int moduli_id) const {
    return kKsExtOffset + 3 * kMaxAlpha * (limbs_at(moduli_id) + alpha);
  }
  std::vector<int64_t> modup_constants(int moduli_id, int key_moduli_id) const;
  // True when the first prime of the partition is at most twice each of its
  // other primes: then the Garner differences of rns_to_mrs
  // (src/ckks/core_key_switch.cpp:21-56) are never negative.
  bool modup_non_negative(int moduli_id, const Partition &part) const {
    const uint64_t first = q[moduli_id + part.start];
    for (int j = 1; j < part.size; j++)
      if (first > 2 * q[moduli_id + part.start + j])
        return false;
    return true;
  }

  // Per-limb scalars of CKKSPrecomputed's multipliers
  // (src/ckks/precomputed.cpp:418-533), limb l of the object at the
  // multiplier's moduli_id:
  //   add_multiplier(value, level)          make_add_multiplier
  //   mult_multiplier(value, level)         make_mult_multiplier(double)
  //   mult_multiplier_int(value, moduli_id) make_mult_multiplier(int64_t)
  //   level_down_multiplier(src, dst)       make_level_down_multiplier
  //   level_down_multiplier_after_rescale   ..._after_rescale
  std::vector<int64_t> add_multiplier(double value, int level) const;
  std::vector<int64_t> mult_multiplier(double value, int level) const;
  std::vector<int64_t> mult_multiplier_int(int64_t value, int moduli_id) const;
  std::vector<int64_t> level_down_multiplier(int source_level,
                                             int destination_level) const;
  std::vector<int64_t>
  level_down_multiplier_after_rescale(int source_level,
                                      int destination_level) const;

  // Encoding tables (src/ckks/precomputed.cpp:50-75, :138-215,
  // src/ckks/encode.cpp:24-150): full slots, codec size 2 * slots = N.
  struct EncodeLevel {
    long double ckks_scale = 0;
    double scale_before_tile = 0;
    int64_t log_scale_after_tile = 0;
    bool top_of_one_more_level = false; // limb 0 zero, values on limbs 1..
    std::vector<int64_t> r_mult_scale_after_tile;
  };
  int slot_count() const { return n / 2; }
  std::vector<uint32_t> encode_pre_permutation() const;
  // cos / sin of index * (-2 pi / codec) (FFT twiddles) and of
  // (2 index) * (-pi / (2 codec)) (the twister), index < codec
  void encode_trig(std::vector<double> &fft_cos, std::vector<double> &fft_sin,
                   std::vector<double> &twist_cos,
                   std::vector<double> &twist_sin) const;
  EncodeLevel encode_level(int level) const;

  // Decoding (src/ckks/core_decode.cpp:10-124): the CRT constants of
  // DecodeParameters (src/parameter/decode.cpp:5-51) for the limbs from
  // moduli_id on, as little-endian 64-bit words.
  struct DecodeCrt {
    int limbs = 0;
    int words = 0;                   // words of Q = prod of the limbs' primes
    std::vector<uint64_t> reverser;  // [limbs][words]: (Q/q_i)((Q/q_i)^-1 mod q_i)
    std::vector<uint64_t> q_product; // Q, `words` words
    std::vector<uint64_t> q_half;    // floor(Q / 2), `words` words
    std::vector<uint64_t> upper;     // Q 2^62, words + 1 words
    uint64_t k = 0;                  // -Q^-1 mod 2^62
  };
  DecodeCrt decode_crt(int moduli_id) const;
  // codec index -> slot (the inverse of the encoding permutation)
  std::vector<uint32_t> decode_post_permutation() const;
  // cos / sin of (2 index) * (pi / (2 codec)) (the skewer) and the inverse
  // FFT twiddles cos(a), sin(-a), a = index * (-2 pi / codec)
  void decode_trig(std::vector<double> &skew_cos, std::vector<double> &skew_sin,
                   std::vector<double> &ifft_cos,
                   std::vector<double> &ifft_sin) const;

private:
  void derive(bool twiddles);
  void derive_deviations();
};

} // namespace fhe
