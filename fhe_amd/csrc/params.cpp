// Host-side CKKS parameter set (see params.h for the reference mapping).
#include "params.h"

#include <algorithm>
#include <cmath>
#include <math.h>
#include <numbers>
#include <stdexcept>

#include "modarith.h"

namespace fhe {

using i128 = __int128;
using u128 = unsigned __int128;

int64_t mod_mul(int64_t a, int64_t b, int64_t m) {
  i128 p = i128(a) * i128(b);
  p %= i128(m);
  return int64_t(p);
}

int64_t mod_pow(int64_t base, int64_t e, int64_t m) {
  int64_t r = 1 % m;
  base %= m;
  for (; e > 0; e >>= 1) {
    if (e & 1)
      r = mod_mul(r, base, m);
    base = mod_mul(base, base, m);
  }
  return r;
}

// Extended Euclid on int64 with the reference's conventions
// (src/arithmetic/arithmetic.cpp:10-31): a result in [0, m), 0 when x has no
// inverse modulo m.
int64_t mod_inverse(int64_t x, int64_t m) {
  int64_t r0 = x, r1 = m, s0 = 1, s1 = 0;
  while (r1 > 0) {
    const int64_t quot = r0 / r1;
    int64_t t = r0 - quot * r1;
    r0 = r1;
    r1 = t;
    t = s0 - quot * s1;
    s0 = s1;
    s1 = t;
  }
  if (r0 != 1)
    return 0;
  return s0 >= 0 ? s0 : s0 + m;
}

static uint64_t pow_u64(uint64_t b, uint64_t e, uint64_t m) {
  u128 r = 1, x = b % m;
  for (; e; e >>= 1) {
    if (e & 1)
      r = (r * x) % m;
    x = (x * x) % m;
  }
  return uint64_t(r);
}

// Deterministic Miller-Rabin for 64-bit integers (the first 12 prime bases
// are a proven witness set below 3.3e24).
bool is_prime_u64(uint64_t n) {
  if (n < 2)
    return false;
  static const uint64_t bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  for (uint64_t p : bases) {
    if (n % p == 0)
      return n == p;
  }
  uint64_t d = n - 1;
  int s = 0;
  while ((d & 1) == 0) {
    d >>= 1;
    s++;
  }
  for (uint64_t a : bases) {
    uint64_t x = pow_u64(a, d, n);
    if (x == 1 || x == n - 1)
      continue;
    bool composite = true;
    for (int i = 1; i < s; i++) {
      x = uint64_t((u128(x) * x) % n);
      if (x == n - 1) {
        composite = false;
        break;
      }
    }
    if (composite)
      return false;
  }
  return true;
}

// Same selection rule as src/arithmetic/arithmetic.cpp:98-113: the first base
// b = 2, 3, ... whose (q-1)/2n-th power has order exactly 2n.
int64_t find_ntt_root(int64_t q, int n) {
  const int64_t unit = (q - 1) / (2 * int64_t(n));
  int64_t e = 0;
  for (int64_t b = 2; b < q; b++) {
    e = mod_pow(b, unit, q);
    if (mod_pow(e, n, q) != 1)
      break;
  }
  return e;
}

// k with R * R^-1 - q * k = 1 (src/arithmetic/arithmetic.cpp:159-167).
uint64_t mont_k(uint64_t q) {
  const int64_t rinv = mod_inverse(kMontR, int64_t(q));
  i128 t = i128(kMontR) * i128(rinv) - 1;
  t /= i128(q);
  return uint64_t(t);
}

static bool is_ntt_prime(uint64_t x, int n) {
  return (x - 1) % (2 * uint64_t(n)) == 0 && is_prime_u64(x);
}

static uint64_t next_ntt_prime(uint64_t start, int n, bool up) {
  uint64_t x = start;
  while (!is_ntt_prime(x, n))
    x = up ? x + 2 : x - 2;
  return x;
}

// src/arithmetic/prime.cpp:913-930: walk down from 2^bits in steps of 2n.
std::vector<uint64_t> decreasing_ntt_primes(int bits, int n, int count) {
  std::vector<uint64_t> out;
  const uint64_t step = 2 * uint64_t(n);
  uint64_t x = ((uint64_t(1) << bits) / step) * step + 1;
  while (int(out.size()) < count) {
    if (is_ntt_prime(x, n))
      out.push_back(x);
    x -= step;
  }
  return out;
}

// Primes alternating above and below 2^bits so the accumulated scale drift
// stays near one (src/arithmetic/prime.cpp:765-843, the multi-precision
// search the reference's cached tables come from).  The drift is tracked in
// binary128 with each prime first rounded to double, which reproduces every
// prime of the reference's cached chains for N = 2^3 .. 2^16.
std::vector<uint64_t> alternating_ntt_primes(int bits, int n, int count) {
  std::vector<uint64_t> out;
  const uint64_t scale = uint64_t(1) << bits;
  uint64_t hi_start = scale + 1, lo_start = scale - 1;
  const uint64_t first_hi = next_ntt_prime(hi_start, n, true);
  const uint64_t first_lo = next_ntt_prime(lo_start, n, false);
  bool up = first_hi - scale >= scale - first_lo;
  __float128 drift = 1;
  const __float128 fscale = __float128(double(scale));
  while (int(out.size()) < count) {
    const uint64_t p = next_ntt_prime(up ? hi_start : lo_start, n, up);
    drift = drift * (fscale / __float128(double(p)));
    drift = drift * drift;
    const uint64_t search = uint64_t(int64_t((drift * fscale) / 2));
    if (up) {
      hi_start = p + 2;
      const uint64_t searched = search * 2 - 1;
      if (searched < lo_start)
        lo_start = searched;
    } else {
      lo_start = p - 2;
      const uint64_t searched = search * 2 + 1;
      if (searched > hi_start)
        hi_start = searched;
    }
    up = !up;
    out.push_back(p);
  }
  return out;
}

// Largest total modulus bit count per ring degree (src/ckks/config.cpp:49-84).
static int max_q_bits(int log_n) {
  switch (log_n) {
  case 3:
  case 4:
    return 1000;
  case 13:
    return 214;
  case 14:
    return 430;
  case 15:
    return 868;
  case 16:
    return 1747;
  case 17:
    return 3523;
  default:
    throw std::invalid_argument("CKKS Parameters: unsupported config.");
  }
}

CkksParams CkksParams::from_config(const CkksConfig &c, bool twiddles) {
  if (c.special_prime_count < 1 || c.special_prime_count > kMaxAlpha)
    throw std::invalid_argument("The given special prime count is invalid.");
  const int n = 1 << c.log_n;
  const int small_count =
      (max_q_bits(c.log_n) - 60 * (1 + c.special_prime_count)) /
      c.small_prime_bits;
  if (small_count < 2)
    throw std::invalid_argument("CKKS Parameters: unsupported config.");
  // src/ckks/parameters.cpp:9-51: small primes, then base + specials, with a
  // copy of large_primes[1] prepended for the one-more-level chain.
  std::vector<uint64_t> primes =
      alternating_ntt_primes(c.small_prime_bits, n, small_count);
  const auto large =
      decreasing_ntt_primes(c.large_prime_bits, n, c.special_prime_count + 1);
  primes.insert(primes.end(), large.begin(), large.end());
  if (c.one_more_level)
    primes.insert(primes.begin(), large[1]);
  CkksParams p = from_primes(c.log_n, primes, c.special_prime_count,
                             c.scale_bits, c.one_more_level, twiddles);
  if (c.max_level >= 0)
    p.max_level = c.max_level;
  return p;
}

CkksParams CkksParams::from_primes(int log_n, const std::vector<uint64_t> &primes,
                                   int special_prime_count, int scale_bits,
                                   bool one_more_level, bool twiddles) {
  CkksParams p;
  p.log_n = log_n;
  p.n = 1 << log_n;
  p.q = primes;
  p.prime_count = int(primes.size());
  p.alpha = special_prime_count;
  p.chain_count = p.prime_count - special_prime_count;
  if (p.chain_count < 2 || special_prime_count < 1 ||
      special_prime_count > kMaxAlpha)
    throw std::invalid_argument("CKKS Parameters: unsupported config.");
  for (auto q : primes)
    if (q >= (uint64_t(1) << 61) || (q - 1) % (2 * uint64_t(p.n)) != 0)
      throw std::invalid_argument("CKKS Parameters: unsupported prime.");
  p.full_level = p.chain_count - 2;
  p.max_level = p.full_level;
  p.scale_bits = scale_bits;
  p.one_more_level = one_more_level;
  p.derive(twiddles);
  return p;
}

int64_t CkksParams::mod_raise_ratio() const {
  const double s = scale;
  const double partial_sum_ratio = 1; // full slots
  // the base prime: the last prime of the ciphertext chain
  const double base = static_cast<double>(q[chain_count - 1]);
  return static_cast<int64_t>(std::round(s * s / (partial_sum_ratio * base)));
}

int CkksParams::max_key_switching_level() const {
  if (one_more_level && max_level == full_level)
    return full_level - 1;
  return max_level;
}

void CkksParams::derive(bool twiddles) {
  derive_deviations();
  const int P = prime_count;
  k.resize(P);
  q2.resize(P);
  qh.resize(P);
  r2.resize(P);
  ninv.resize(P);
  root.assign(twiddles ? size_t(P) * n : 0, 0);
  root_inverse.assign(twiddles ? size_t(P) * n : 0, 0);
  root_shoup.assign(twiddles ? size_t(P) * n * 2 : 0, 0);
  root_inverse_shoup.assign(twiddles ? size_t(P) * n * 2 : 0, 0);
  std::vector<int64_t> power(n), shuffled(n);
  auto bitrev = [&](int x) {
    int y = 0;
    for (int b = 0; b < log_n; b++)
      y |= ((x >> b) & 1) << (log_n - 1 - b);
    return y;
  };
  for (int i = 0; i < P; i++) {
    const int64_t qi = int64_t(q[i]);
    q2[i] = q[i] << 1;
    qh[i] = q[i] >> 1;
    k[i] = mont_k(q[i]);
    r2[i] = mod_mul(kMontR, kMontR, qi);
    ninv[i] = mod_mul(mod_inverse(n, qi), kMontR, qi);
    if (!twiddles)
      continue;
    const int64_t psi = find_ntt_root(qi, n);
    for (int pass = 0; pass < 2; pass++) {
      const int64_t base = pass == 0 ? psi : mod_inverse(psi, qi);
      power[0] = 1;
      for (int j = 1; j < n; j++)
        power[j] = mod_mul(power[j - 1], base, qi);
      for (int j = 0; j < n; j++)
        shuffled[bitrev(j)] = power[j];
      int64_t *dst = (pass == 0 ? root.data() : root_inverse.data()) +
                     size_t(i) * n;
      if (pass == 0) {
        // forward: bit-reversed powers 1..N-1 (CT stage s, group g uses
        // entry 2^s - 1 + g)
        for (int j = 1; j < n; j++)
          dst[j - 1] = shuffled[j];
      } else {
        // inverse: stage t reads the N/2^(t+1) entries starting at
        // N - N/2^t (src/ckks/precomputed.cpp:15-31)
        int o = 0;
        for (int lg = log_n - 1; lg >= 0; lg--)
          for (int g = 0; g < (1 << lg); g++)
            dst[o++] = shuffled[(1 << lg) + g];
      }
      uint64_t *sh = (pass == 0 ? root_shoup.data() : root_inverse_shoup.data()) +
                     size_t(i) * n * 2;
      for (int j = 0; j < n - 1; j++) {
        sh[2 * j] = uint64_t(dst[j]);
        sh[2 * j + 1] = uint64_t(((unsigned __int128)dst[j] << 64) / q[i]);
        dst[j] = mont_mult_u(dst[j], r2[i], q[i], k[i]);
      }
    }
  }
  // src/ckks/precomputed.cpp:97-137
  p_product_mult_r.assign(chain_count, 0);
  for (int i = 0; i < chain_count; i++) {
    int64_t prod = kMontR;
    for (int s = chain_count; s < P; s++)
      prod = mod_mul(prod, int64_t(q[s]), int64_t(q[i]));
    p_product_mult_r[i] = prod;
  }
  q_top.resize(P);
  for (int i = 0; i < P; i++)
    q_top[i] = ((uint64_t(1) << 63) - 1) / q[i] * q[i];
  p_inverse_mult_r.assign(size_t(alpha) * P, 0);
  p_inverse_shoup.assign(size_t(alpha) * P * 2, 0);
  for (int sp = 0; sp < alpha; sp++) {
    const int pi = P - 1 - sp;
    for (int i = 0; i < pi; i++) {
      const int64_t inv = mod_inverse(int64_t(q[pi]), int64_t(q[i]));
      p_inverse_mult_r[size_t(sp) * P + i] =
          mod_mul(inv, kMontR, int64_t(q[i]));
      p_inverse_shoup[(size_t(sp) * P + i) * 2] = uint64_t(inv);
      p_inverse_shoup[(size_t(sp) * P + i) * 2 + 1] =
          uint64_t(((unsigned __int128)uint64_t(inv) << 64) / q[i]);
    }
  }
  rescale_scale.assign(size_t(P) * P, 0);
  for (int from = 0; from < P; from++)
    for (int to = from + 1; to < P; to++)
      rescale_scale[size_t(from) * P + to] =
          mod_mul(mod_inverse(int64_t(q[from]), int64_t(q[to])), kMontR,
                  int64_t(q[to]));
}

std::vector<Partition> CkksParams::partitions(int moduli_id,
                                              int key_moduli_id) const {
  const int chain = limbs_at(moduli_id);
  const int count = partition_count(moduli_id);
  int key_index = partition_count(key_moduli_id) - count;
  std::vector<Partition> out;
  for (int start = chain - count * alpha; start < chain;
       start += alpha, key_index++) {
    const int lo = std::max(0, start);
    const int hi = std::min(chain, start + alpha);
    out.push_back({lo, hi - lo, key_index});
  }
  return out;
}

// (prod_{t<j} q[moduli_id + start + t]) mod m
static int64_t prefix_product_mod(const CkksParams &p, int moduli_id, int start,
                                  int j, int64_t m) {
  int64_t r = 1 % m;
  for (int t = 0; t < j; t++)
    r = mod_mul(r, int64_t(p.q[moduli_id + start + t] % uint64_t(m)), m);
  return r;
}

std::vector<int64_t> CkksParams::modup_constants(int moduli_id,
                                                 int key_moduli_id) const {
  const auto parts = partitions(moduli_id, key_moduli_id);
  const int ext = limbs_at(moduli_id) + alpha;
  const int stride = ks_stride(moduli_id);
  std::vector<int64_t> c(size_t(stride) * parts.size(), 0);
  for (size_t pi = 0; pi < parts.size(); pi++) {
    const auto &part = parts[pi];
    int64_t *blk = c.data() + pi * stride;
    for (int j = 1; j < part.size; j++) {
      const int64_t m = int64_t(q[moduli_id + part.start + j]);
      const int64_t prod = prefix_product_mod(*this, moduli_id, part.start, j, m);
      blk[kKsInvOffset + j] = mod_mul(mod_inverse(prod, m), kMontR, m);
      for (int t = j + 1; t < part.size; t++) {
        const int64_t mt = int64_t(q[moduli_id + part.start + t]);
        blk[kKsTileOffset + j * kMaxAlpha + t] = mod_mul(
            prefix_product_mod(*this, moduli_id, part.start, j, mt), kMontR, mt);
      }
    }
    const int shoup = ks_ext_shoup_offset(moduli_id);
    for (int j = 0; j < part.size; j++)
      for (int i = 0; i < ext; i++) {
        const int64_t m = int64_t(q[moduli_id + i]);
        const int64_t prod =
            prefix_product_mod(*this, moduli_id, part.start, j, m);
        blk[kKsExtOffset + j * ext + i] =
            j == 0 ? r2[moduli_id + i] : mod_mul(prod, r2[moduli_id + i], m);
        const uint64_t w = uint64_t(mod_mul(prod, kMontR, m));
        blk[shoup + 2 * (j * ext + i)] = int64_t(w);
        blk[shoup + 2 * (j * ext + i) + 1] =
            int64_t(uint64_t(((unsigned __int128)w << 64) / uint64_t(m)));
      }
  }
  return c;
}


// ---------------------------------------------------------------- scales

namespace {

// floor(x) mod q for a finite x >= 0, where floor(x) is the integer the
// reference builds digit by digit in base 2^64 (BigInt::from_double,
// src/arithmetic/big_int.cpp:29-42): each digit is the double remainder
// truncated to an integer.
uint64_t floor_double_mod(double x, uint64_t q) {
  const double base = std::pow(2, 64);
  std::vector<unsigned __int128> digits;
  while (x > 0) {
    const double quotient = std::floor(x / base);
    const double remainder = x - quotient * base;
    x = quotient;
    digits.push_back(static_cast<unsigned __int128>(remainder));
  }
  unsigned __int128 acc = 0;
  const unsigned __int128 base_mod = ((unsigned __int128)1 << 64) % q;
  for (size_t i = digits.size(); i-- > 0;)
    acc = (acc * base_mod + digits[i] % q) % q;
  return uint64_t(acc);
}

// scale_modulo / scale_enter_modulo (src/arithmetic/arithmetic.cpp:60-95):
// the double `product` (operation order as there), its floor reduced per
// prime, times R when entering, negated as q - s for a negative product.
std::vector<int64_t> signed_scalar(double product, bool enter,
                                   const uint64_t *q, int count) {
  std::vector<int64_t> out(count);
  for (int i = 0; i < count; i++) {
    uint64_t s = floor_double_mod(std::abs(product), q[i]);
    if (enter)
      s = uint64_t((unsigned __int128)s * (uint64_t(kMontR) % q[i]) % q[i]);
    int64_t v = int64_t(s);
    if (product < 0)
      v = int64_t(q[i]) - v;
    out[i] = v;
  }
  return out;
}

} // namespace

void CkksParams::derive_deviations() {
  scale = std::pow(2, scale_bits);
  // levels full_level .. 0 in order, as src/ckks/parameters.cpp:88-97
  std::vector<long double> dev;
  for (int level = full_level, moduli_id = 0; level >= 0; level--, moduli_id++) {
    if (level == full_level || (one_more_level && level == full_level - 1)) {
      dev.push_back(1.0);
    } else {
      const long double level_q = static_cast<long double>(q[moduli_id - 1]);
      const auto alpha_sq = std::pow(scale / level_q, 2);
      dev.push_back(std::pow(dev.back(), 2) * alpha_sq);
    }
  }
  deviation_ld.assign(dev.rbegin(), dev.rend());
  deviation.resize(deviation_ld.size());
  for (size_t i = 0; i < deviation_ld.size(); i++)
    deviation[i] = static_cast<double>(deviation_ld[i]);
}

std::vector<int64_t> CkksParams::add_multiplier(double value, int level) const {
  const int mid = moduli_id_of_level(level);
  double dev = deviation.at(level);
  if (one_more_level && mid == 0)
    dev = q[0] / scale;
  const double product = (value * scale * dev + 0.5) * scale;
  return signed_scalar(product, false, q.data() + mid, limbs_at(mid));
}

std::vector<int64_t> CkksParams::mult_multiplier(double value,
                                                 int level) const {
  const int mid = moduli_id_of_level(level);
  const double dev = deviation.at(level - 1);
  const double product = value * scale * std::sqrt(dev) + 0.5;
  return signed_scalar(product, true, q.data() + mid, limbs_at(mid));
}

std::vector<int64_t> CkksParams::mult_multiplier_int(int64_t value,
                                                     int moduli_id) const {
  // modulus_multiply (src/arithmetic/arithmetic.cpp:33-38): a truncated
  // 128-bit remainder, negative for a negative value
  std::vector<int64_t> out(limbs_at(moduli_id));
  for (size_t i = 0; i < out.size(); i++) {
    const __int128 prod = (__int128)value * (__int128)kMontR;
    out[i] = int64_t(prod % (__int128)q[moduli_id + i]);
  }
  return out;
}

std::vector<int64_t>
CkksParams::level_down_multiplier(int source_level,
                                  int destination_level) const {
  const int src_mid = moduli_id_of_level(source_level);
  const double diff = deviation.at(destination_level) /
                      deviation.at(source_level);
  double delta;
  if (one_more_level && src_mid == 0)
    delta = scale * diff;
  else
    delta = q[src_mid] * diff;
  return signed_scalar(std::round(delta), true, q.data() + src_mid,
                       limbs_at(src_mid));
}

std::vector<int64_t>
CkksParams::level_down_multiplier_after_rescale(int source_level,
                                                int destination_level) const {
  const double diff = deviation.at(destination_level) /
                      std::sqrt(deviation.at(source_level - 1));
  const int dst_mid = moduli_id_of_level(destination_level);
  return signed_scalar(std::round(scale * diff), true, q.data() + dst_mid,
                       limbs_at(dst_mid));
}

// ---------------------------------------------------------------- encoding

namespace {

// The slot ordering of the canonical embedding: the circular shift of the
// two halves conjugated by the powers-of-5 fold, inverted
// (src/ckks/encode.cpp:24-150 with the steps of src/ckks/precomputed.cpp:
// 50-58).  A permutation conjugation maps the k-th element of the cycle
// listing of the first onto the k-th element of the second's listing; the
// listing visits cycles from their smallest unvisited start, each from its
// start's image.
std::vector<uint64_t> cycle_listing(const std::vector<uint64_t> &perm) {
  const size_t n = perm.size();
  std::vector<char> seen(n, 0);
  std::vector<uint64_t> out;
  out.reserve(n);
  size_t start = 0;
  while (start < n) {
    uint64_t cur = perm[start];
    uint64_t next = perm[cur];
    while (true) {
      out.push_back(cur);
      seen[cur] = 1;
      cur = next;
      if (seen[next])
        break;
      next = perm[cur];
    }
    start++;
    while (start < n && seen[start])
      start++;
  }
  return out;
}

} // namespace

std::vector<uint32_t> CkksParams::encode_pre_permutation() const {
  const int size = n; // codec coefficient count for full slots
  const int half = size / 2;
  std::vector<uint64_t> shift(size), fold(size);
  shift[0] = half - 1;
  for (int i = 1; i < half; i++)
    shift[i] = i - 1;
  for (int i = half; i < size - 1; i++)
    shift[i] = i + 1;
  shift[size - 1] = half;
  // fold: (g (2 i + 1) mod 2 size - 1) / 2 with g = 5^(2 size - 1), the
  // inverse of 5 modulo 2 size (src/ckks/encode.cpp:44-87), as a running sum
  const int64_t two = 2 * int64_t(size);
  int64_t jump = 1;
  for (int64_t e = 0; e < two - 1; e++)
    jump = (jump * 5) & (two - 1);
  int64_t acc = jump;
  for (int i = 0; i < size; i++) {
    fold[i] = uint64_t((acc - 1) >> 1);
    acc = (acc + jump * 2) % two;
  }
  const auto c1 = cycle_listing(fold), c2 = cycle_listing(shift);
  std::vector<uint64_t> post(size);
  for (int i = 0; i < size; i++)
    post[c1[i]] = c2[i];
  std::vector<uint32_t> pre(size);
  for (int i = 0; i < size; i++)
    pre[post[i]] = uint32_t(i);
  return pre;
}

namespace {

// libm's sin and sincos, called through volatile pointers so that the
// compiler neither folds nor pairs them: the reference's -O2 build computes
// every same-argument (cos x, sin x) pair with glibc's sincos, whose results
// differ from separate sin and cos calls in the last bit for some x, and
// computes sin(-x) next to such a pair with a separate sin call.
double (*volatile libm_sin)(double) = ::sin;
void (*volatile libm_sincos)(double, double *, double *) = ::sincos;

} // namespace

void CkksParams::encode_trig(std::vector<double> &fft_cos,
                             std::vector<double> &fft_sin,
                             std::vector<double> &twist_cos,
                             std::vector<double> &twist_sin) const {
  const int codec = n;
  const double pi = std::numbers::pi;
  fft_cos.resize(codec);
  fft_sin.resize(codec);
  twist_cos.resize(codec);
  twist_sin.resize(codec);
  const double unit_angle = -2 * pi / codec;
  const double step = -pi / (codec * 2);
  for (int i = 0; i < codec; i++) {
    const double angle = i * unit_angle;
    libm_sincos(angle, &fft_sin[i], &fft_cos[i]);
    const int real_index = 2 * i;
    libm_sincos(real_index * step, &twist_sin[i], &twist_cos[i]);
  }
}

CkksParams::EncodeLevel CkksParams::encode_level(int level) const {
  EncodeLevel e;
  int mid = moduli_id_of_level(level);
  int chain = limbs_at(mid) + alpha;
  const int log_scale = scale_bits;
  const long double sc = scale;
  e.ckks_scale = static_cast<long double>(deviation.at(level)) * sc * sc;
  if (one_more_level && mid == 0)
    e.ckks_scale = sc * static_cast<long double>(q[0]);
  const uint64_t p0 = q[mid];
  e.top_of_one_more_level =
      one_more_level && mid == 0 && (e.ckks_scale >= p0 * scale);
  if (e.top_of_one_more_level) {
    e.scale_before_tile = double(e.ckks_scale / p0);
    chain -= 1;
    mid = moduli_id_of_level(level - 1);
  } else {
    int exponent;
    const double mantissa = double(std::frexp(e.ckks_scale, &exponent));
    if (exponent >= log_scale) {
      e.scale_before_tile = mantissa * scale;
      e.log_scale_after_tile = exponent - log_scale;
    } else {
      e.scale_before_tile = double(e.ckks_scale);
      e.log_scale_after_tile = 0;
    }
  }
  e.r_mult_scale_after_tile.resize(chain);
  for (int i = 0; i < chain; i++) {
    const int64_t qt = int64_t(q[mid + i]);
    if (e.top_of_one_more_level) {
      e.r_mult_scale_after_tile[i] = int64_t(
          (__int128)kMontR * (__int128)int64_t(p0) % (__int128)qt);
    } else {
      int64_t v = kMontR % qt;
      for (int64_t ls = e.log_scale_after_tile; ls > 0; ls -= 60) {
        const int64_t part = int64_t(1) << std::min<int64_t>(60, ls);
        v = int64_t((__int128)v * part % qt);
      }
      e.r_mult_scale_after_tile[i] = v;
    }
  }
  return e;
}

namespace {

// Little-endian multiword helpers for the decoding constants.
using Words = std::vector<uint64_t>;

void trim(Words &a) {
  while (a.size() > 1 && a.back() == 0)
    a.pop_back();
}

Words mul_word(const Words &a, uint64_t b) {
  Words out(a.size() + 1, 0);
  u128 carry = 0;
  for (size_t i = 0; i < a.size(); i++) {
    const u128 p = u128(a[i]) * b + carry;
    out[i] = uint64_t(p);
    carry = p >> 64;
  }
  out[a.size()] = uint64_t(carry);
  trim(out);
  return out;
}

// exact quotient a / b and the remainder
Words div_word(const Words &a, uint64_t b, uint64_t *rem) {
  Words out(a.size(), 0);
  u128 r = 0;
  for (size_t i = a.size(); i-- > 0;) {
    const u128 cur = (r << 64) | a[i];
    out[i] = uint64_t(cur / b);
    r = cur % b;
  }
  if (rem)
    *rem = uint64_t(r);
  trim(out);
  return out;
}

} // namespace

CkksParams::DecodeCrt CkksParams::decode_crt(int moduli_id) const {
  DecodeCrt d;
  d.limbs = limbs_at(moduli_id);
  Words prod{1};
  for (int i = 0; i < d.limbs; i++)
    prod = mul_word(prod, q[moduli_id + i]);
  d.words = int(prod.size());
  auto pad = [&](Words w, int n) {
    w.resize(size_t(n), 0);
    return w;
  };
  d.q_product = pad(prod, d.words);
  uint64_t rem = 0;
  d.q_half = pad(div_word(prod, 2, &rem), d.words);
  d.upper = pad(mul_word(prod, uint64_t(1) << 62), d.words + 1);
  d.reverser.assign(size_t(d.limbs) * d.words, 0);
  for (int i = 0; i < d.limbs; i++) {
    const uint64_t qi = q[moduli_id + i];
    const Words cofactor = div_word(prod, qi, nullptr);
    uint64_t cof_mod = 0;
    div_word(cofactor, qi, &cof_mod);
    const int64_t inv = mod_inverse(int64_t(cof_mod), int64_t(qi));
    const Words r = pad(mul_word(cofactor, uint64_t(inv)), d.words);
    std::copy(r.begin(), r.end(), d.reverser.begin() + size_t(i) * d.words);
  }
  // -Q^-1 mod 2^62 from Q mod 2^64 (Newton iteration for the inverse)
  const uint64_t q0 = prod[0];
  uint64_t inv = q0; // correct to 3 bits for odd q0
  for (int it = 0; it < 6; it++)
    inv *= 2 - q0 * inv;
  d.k = (uint64_t(0) - inv) & ((uint64_t(1) << 62) - 1);
  return d;
}

std::vector<uint32_t> CkksParams::decode_post_permutation() const {
  const auto pre = encode_pre_permutation();
  std::vector<uint32_t> post(pre.size());
  for (size_t i = 0; i < pre.size(); i++)
    post[pre[i]] = uint32_t(i);
  return post;
}

void CkksParams::decode_trig(std::vector<double> &skew_cos,
                             std::vector<double> &skew_sin,
                             std::vector<double> &ifft_cos,
                             std::vector<double> &ifft_sin) const {
  const int codec = n;
  const double pi = std::numbers::pi;
  skew_cos.resize(codec);
  skew_sin.resize(codec);
  ifft_cos.resize(codec);
  ifft_sin.resize(codec);
  const double step = pi / (codec * 2);
  const double unit_angle = -2 * pi / codec;
  for (int i = 0; i < codec; i++) {
    const int real_index = 2 * i;
    libm_sincos(real_index * step, &skew_sin[i], &skew_cos[i]);
    const double angle = i * unit_angle;
    double unused;
    libm_sincos(angle, &unused, &ifft_cos[i]);
    ifft_sin[i] = libm_sin(-angle);
  }
}

} // namespace fhe
