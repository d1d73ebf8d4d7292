// Engine-level compositions on the device (CKKSEngine's form bookkeeping,
// EvalMod, the slot transforms and the bootstrap), host C++ over the device
// operations of the C ABI.
#include <set>

#include "text_engine.h"

namespace {

// The Chebyshev coefficients of the reference's sine approximation
// (src/ckks/engine_bootstrap.cpp:500-524): part of the algorithm's
// definition, so the same values.
const std::vector<double> kSineAuxiliary = {
    -0.4674215360340989,   -0.24807246365084598, -0.03360736343811264,
    0.10485552200765141,   0.012171628025753911, -0.031034212401993305,
    -0.003985828764410408, 0.0};
const std::vector<std::vector<double>> kSineQuotient = {
    {-0.7500012154934053, 4.148627473785995e-06, 2.707983311367457e-07,
     -4.324538856989888e-07, -2.648274421485692e-08, 3.9515438249932484e-08,
     2.4720260690299275e-09, 0.0},
    {-6.511205244110998e-09, -7.076428529570776e-10, 1.0183620860162692e-09,
     0.0, 0.0, 0.0, 0.0, 4.0},
    {0.007868944223231958, 0.0014193060148561195, -0.003081157124604844,
     -0.00025275745577540515, 0.0005030348400875631, 3.734766164626263e-05,
     -7.780382338014487e-05, 1.0}};
const std::vector<std::vector<double>> kSineRemainder = {
    {-1.1324455480972702, -0.8116709016985189, -0.20294197180913795,
     0.5290215620326779, -0.32084902738708476, 0.43458317383580536,
     -0.06303880005650665, 0.0},
    {0.23316927298939902, 0.12938749185636858, -0.3042902283756468,
     -0.008967445477730096, -0.6955791340472846, -0.07562064123442451,
     1.2510854562849574, 2.0},
    {0.1429148264191405, 0.4544397978460022, -0.8401926425765848,
     -0.4292803758141986, -1.495563076527614, 0.3580084476476745,
     1.7971594278092713, 1.0}};
constexpr int kBabyDegree = 7;
constexpr int kGiantDegree = 3;
constexpr int kDoubleAngleIterations = 3;
constexpr int64_t kRoundBound = 28;
// the double nearest pi (std::numbers::pi in the reference)
constexpr double kPi = 3.141592653589793238462643383279502884;

// chebyshev_paterson_stockmeyer (src/ckks/engine_bootstrap.cpp:459-498)
Text chebyshev_ps(TextEngine &e, const std::vector<Text> &baby,
                  const Text &giant,
                  const std::vector<std::vector<double>> &c, bool relin) {
  Text baby0 = e.weighted_sum(baby, c[0]);
  Text added = e.add(giant, baby0, false);
  Text baby1 = e.weighted_sum(baby, c[1]);
  Text multiplied = e.multiply(added, baby1, relin);
  Text baby2 = e.weighted_sum(baby, c[2]);
  return e.add(multiplied, baby2, false);
}

// sine (src/ckks/engine_bootstrap.cpp:526-638): Chebyshev basis T1..T7,
// giant steps T7, T14, T28 and the Paterson-Stockmeyer evaluation
Text sine(TextEngine &e, const Text &x) {
  // texts are shared, not copied: the only in-place additions below are
  // on fresh products
  std::vector<Text> baby(kBabyDegree);
  baby[0] = x;
  for (int bi = 2; bi <= kBabyDegree; bi++) {
    const int index = bi - 1, half = bi / 2 - 1;
    if ((bi & 1) == 0) {
      Text m = e.multiply(baby[half], baby[half], true);
      e.add_inplace(m, m);
      baby[index] = e.add_double(m, -1);
    } else {
      Text m = e.multiply(baby[half], baby[half + 1], true);
      e.add_inplace(m, m);
      baby[index] = e.add(m, baby[0], true);
    }
  }
  const int destination = baby.back().level;
  for (int i = 0; i < kBabyDegree - 2; i++)
    if (baby[i].level > destination)
      baby[i] = e.level_down(baby[i], destination);

  std::vector<Text> giant(kGiantDegree);
  giant[0] = baby.back();
  for (int gi = 1; gi < kGiantDegree; gi++) {
    Text m = e.multiply(giant[gi - 1], giant[gi - 1], true);
    e.add_inplace(m, m);
    giant[gi] = e.add_double(m, -1);
  }
  Text ps = giant[0];
  for (int gi = 1; gi < kGiantDegree - 1; gi++) {
    Text m = e.multiply(ps, giant[gi], true);
    e.add_inplace(m, m);
    ps = e.add(m, giant[0], true);
  }
  {
    Text m = e.multiply(ps, giant[kGiantDegree - 1], false);
    e.add_inplace(m, m);
    ps = e.add(m, giant[0], true);
  }
  Text aux = e.weighted_sum(baby, kSineAuxiliary);
  Text added = e.add(aux, giant.back(), false);
  Text quotient =
      chebyshev_ps(e, baby, giant[kGiantDegree - 2], kSineQuotient, true);
  Text multiplied = e.multiply(added, quotient, false);
  Text remainder =
      chebyshev_ps(e, baby, giant[kGiantDegree - 2], kSineRemainder, false);
  added = e.add(multiplied, remainder, false);
  Text subtracted = e.add(added, ps, true);
  return e.relinearize(subtracted);
}

// apply_double_angle_iterations (src/ckks/engine_bootstrap.cpp:640-659)
Text double_angle(TextEngine &e, const Text &x) {
  Text current = x; // shared: the loop replaces it, never adds into it
  const double two_pi = 2 * kPi;
  for (int it = 1; it <= kDoubleAngleIterations; it++) {
    Text m = e.multiply(current, current, true);
    e.add_inplace(m, m);
    const double scalar =
        -std::pow(two_pi, -std::pow(2, it - kDoubleAngleIterations));
    current = e.add_double(m, scalar);
  }
  return current;
}

using Complex = std::complex<double>;
using Vectors = std::vector<std::vector<Complex>>;

int positive_modulo(int x, int m) { return ((x % m) + m) % m; }

// glibc's sincos through a volatile pointer (neither folded nor split into
// sin and cos, whose results differ from it in the last bit for some x)
void (*volatile libm_sincos_fn)(double, double *, double *) = ::sincos;

// get_phi / get_inverse_phi (src/operation/bootstrap.cpp:9-33): the powers
// of exp(-+ 2 pi i / n).  The reference's complex exponential of (0, y) is
// glibc's cexp: (cos y, sin y) from one sincos call (exp(0) = 1 leaves them
// unchanged), y itself as the sine when |y| is not above DBL_MIN.
std::vector<Complex> phi_table(int n, bool inverse) {
  std::vector<Complex> out(n);
  const double pi_over = kPi / n;
  for (int i = 0; i < n; i++) {
    const double y = inverse ? 2 * pi_over * i : -2 * pi_over * i;
    double sn = y, cs = 1;
    if (std::fabs(y) > DBL_MIN)
      libm_sincos_fn(y, &sn, &cs);
    out[i] = Complex(cs, sn);
  }
  return out;
}

struct Twiddles {
  Vectors center, left, right;
};

// get_twiddle_factor / get_inverse_twiddle_factor
// (src/operation/bootstrap.cpp:35-123): the per-level butterfly diagonals
// of the slot transform in the rotation-by-5 ordering
Twiddles twiddle_factors(int slots, int log_slots, bool inverse) {
  Twiddles t;
  const auto phi = phi_table(slots << 2, inverse);
  int jump_size = inverse ? 1 : slots;
  for (int step = 0; step < log_slots; step++) {
    const int log_index = inverse ? log_slots - 1 - step : step;
    const int jump_count = 1 << log_index;
    if (!inverse)
      jump_size >>= 1;
    std::vector<int> phi_index;
    int power_of_five = 1;
    for (int i = 0; i < jump_size; i++) {
      phi_index.push_back(power_of_five * jump_count);
      power_of_five *= 5;
      power_of_five &= (jump_size << 3) - 1;
    }
    std::vector<Complex> center(slots, 0), left(slots, 0), right(slots, 0);
    for (int j = 0; j < jump_count; j++) {
      int index = 2 * j * jump_size;
      for (int i = 0; i < jump_size; i++) {
        const Complex value = phi[phi_index[i]];
        center[index] = 1;
        center[jump_size + index] = -value;
        if (inverse) {
          left[index] = value;
          right[jump_size + index] = 1;
        } else {
          left[index] = 1;
          right[jump_size + index] = value;
        }
        index++;
      }
    }
    t.center.push_back(center);
    t.left.push_back(left);
    t.right.push_back(right);
    if (inverse)
      jump_size <<= 1;
  }
  return t;
}

// get_{coeff_to_slot,slot_to_coeff}_log_radix_count
// (src/operation/bootstrap.h:23-39)
int cts_log_radix(int stage, int stages, int log_slots) {
  return stage >= stages - log_slots % stages ? log_slots / stages + 1
                                              : log_slots / stages;
}
int stc_log_radix(int stage, int stages, int log_slots) {
  return stage < log_slots % stages ? log_slots / stages + 1
                                    : log_slots / stages;
}

// get_coeff_to_slot_twiddles / get_slot_to_coeff_twiddles
// (src/operation/bootstrap.cpp:125-263): the merged diagonals of each
// stage's radix-2^r butterflies, the last CtS stage normalized by
// 2 * slots * 28
std::vector<Vectors> stage_twiddles(int slots, int log_slots, int stages,
                                    bool to_slot) {
  const Twiddles f = twiddle_factors(slots, log_slots, !to_slot);
  std::vector<Vectors> out;
  int radix_jump = to_slot ? slots : 1;
  int offset = 0;
  for (int stage = 0; stage < stages; stage++) {
    const int log_radix = to_slot ? cts_log_radix(stage, stages, log_slots)
                                  : stc_log_radix(stage, stages, log_slots);
    const int radix = 1 << log_radix;
    const int total = 2 * radix - 1;
    Vectors st(total, std::vector<Complex>(slots, 1));
    for (int lr = 0; lr < log_radix; lr++) {
      if (to_slot)
        radix_jump >>= 1;
      const int radix_index = 1 << lr;
      const int jump = to_slot ? radix / (2 * radix_index) : radix_index;
      const int ti = offset + lr;
      Vectors prev(total, std::vector<Complex>(slots, 0));
      st.swap(prev);
      const auto &center = f.center[ti];
      const auto &left = f.left[ti];
      const auto &right = f.right[ti];
      for (int index = -radix_index + 1; index < radix_index; index++) {
        const int unit = to_slot ? index * radix / radix_index : index;
        const int ci = positive_modulo(unit, total);
        const int li = positive_modulo(unit - jump, total);
        const int ri = positive_modulo(unit + jump, total);
        for (int sl = 0; sl < slots; sl++) {
          st[ci][sl] += center[sl] * prev[ci][sl];
          st[li][sl] +=
              left[sl] * prev[ci][positive_modulo(sl + radix_jump, slots)];
          st[ri][sl] +=
              right[sl] * prev[ci][positive_modulo(sl - radix_jump, slots)];
        }
      }
      if (!to_slot)
        radix_jump <<= 1;
    }
    if (to_slot && stage == stages - 1) {
      const int normalizing = 2 * slots * int(kRoundBound);
      for (auto &v : st)
        for (auto &c : v)
          c /= normalizing;
    }
    out.push_back(std::move(st));
    offset += log_radix;
  }
  return out;
}

// get_stage_loop_counts (src/operation/bootstrap.cpp:265-278)
std::pair<int, int> stage_loop_counts(int twiddle_count) {
  int bits = 0;
  for (int t = twiddle_count; t > 0; t >>= 1)
    bits++;
  const int inner = 1 << ((bits + 1) / 2);
  return {inner, (twiddle_count + 1) / inner};
}

// get_stage_loop_counts with the BootstrapKey's inner-loop shift
// (src/operation/bootstrap.cpp:265-278)
std::pair<int, int> stage_loop_counts_shifted(int twiddle_count, int shift) {
  int bits = 0;
  for (int t = twiddle_count; t > 0; t >>= 1)
    bits++;
  const int inner = 1 << ((bits + 1) / 2 + shift);
  return {inner, (twiddle_count + 1) / inner};
}

// The stage counts a BootstrapKey allows (require_bootstrap_stage_count,
// src/ckks/engine_require.cpp:1018-1100): a table of the reference's.
bool bootstrap_stage_count_valid(int log_slots, int shift, int stages) {
  int lo = 3, hi = 5;
  if (shift == 0) {
    if (log_slots <= 3)
      lo = hi = 1;
    else if (log_slots <= 5)
      lo = 1, hi = 2;
    else if (log_slots == 6)
      lo = hi = 2;
    else if (log_slots <= 9)
      lo = 2, hi = 3;
    else if (log_slots == 10)
      lo = 2, hi = 4;
    else if (log_slots <= 12)
      lo = 3, hi = 4;
    else if (log_slots <= 14)
      lo = 3, hi = 5;
  } else if (shift == 1) {
    if (log_slots <= 2)
      return false;
    if (log_slots <= 5)
      lo = hi = 1;
    else if (log_slots <= 8)
      lo = hi = 2;
    else if (log_slots <= 10)
      lo = 2, hi = 3;
    else if (log_slots == 11)
      lo = hi = 3;
    else if (log_slots <= 14)
      lo = 3, hi = 4;
  } else {
    return false;
  }
  return stages >= lo && stages <= hi;
}

struct StageLayout {
  int level = 0, inner = 0, outer = 0;
  std::vector<int> inner_deltas, outer_deltas;
};

// One PlainMatrix's shape as make_coeff_to_slot_stages /
// make_slot_to_coeff_stages and make_bootstrap_stage build it
// (src/ckks/engine_bootstrap.cpp:8-111, stage levels :113-123)
StageLayout stage_layout(const fhe::CkksParams &P, int stages, int shift,
                         bool to_slot, int stage) {
  const int slots = int(P.slot_count());
  int log_slots = 0;
  while ((1 << log_slots) < slots)
    log_slots++;
  const int top = P.max_key_switching_level();
  StageLayout s;
  int rotate_count = to_slot ? slots : 1;
  for (int st = 0; st <= stage; st++) {
    if (to_slot)
      rotate_count >>= cts_log_radix(st, stages, log_slots);
    else if (st < stage)
      rotate_count <<= stc_log_radix(st, stages, log_slots);
  }
  const int log_radix = to_slot ? cts_log_radix(stage, stages, log_slots)
                                : stc_log_radix(stage, stages, log_slots);
  s.level = to_slot ? top - stage : top - stages - 9 - stage;
  const auto [inner, outer] =
      stage_loop_counts_shifted(2 * (1 << log_radix) - 1, shift);
  s.inner = inner;
  s.outer = outer;
  s.inner_deltas.push_back(0);
  for (int i = 1; i < inner; i++)
    s.inner_deltas.push_back(positive_modulo(i * rotate_count, slots));
  for (int o = 0; o < outer; o++)
    s.outer_deltas.push_back(
        positive_modulo(inner * (o - outer / 2) * rotate_count, slots));
  return s;
}

// rotate_vector (src/ckks/engine_bootstrap.cpp:856-870)
std::vector<Complex> rotate_vector(const std::vector<Complex> &in,
                                   int delta) {
  const int count = int(in.size());
  const int d = (delta + 2 * count) % count;
  std::vector<Complex> out(count);
  for (int i = 0; i < count; i++)
    out[(i + d) % count] = in[i];
  return out;
}

// apply_butterfly with a RotationKey (src/ckks/engine_bootstrap.cpp:
// 155-215): baby-step rotations, diagonal products, giant-step rotations
Text apply_butterfly(TextEngine &e, const Text &x, const Vectors &tw,
                     int rotate_count, bool to_slot, int stages, int stage) {
  const int count = int(tw.size());
  const auto [inner, outer] = stage_loop_counts(count);
  std::vector<int> deltas;
  for (int i = 0; i < inner; i++)
    deltas.push_back(rotate_count * i);
  const std::vector<Text> rotated = e.rotate_batch(x, deltas);
  // every (outer, inner) diagonal, rotated as the reference rotates it,
  // encoded in one batch at the rotated texts' level and prepared; kept
  // resident in the context for later bootstraps
  const int level = rotated[0].level;
  fhe_context *ctx = e.context();
  const auto key = std::make_tuple(to_slot, stages, stage, level);
  auto it = ctx->boot_diagonals.find(key);
  if (it == ctx->boot_diagonals.end()) {
    const int slots = int(tw[0].size());
    std::vector<Complex> factors;
    for (int o = 0; o < outer; o++) {
      const int nidx = o - outer / 2;
      const int tf_delta = inner * nidx * rotate_count;
      for (int i = (o == 0 ? 1 : 0); i < inner; i++) {
        const int ti = inner * nidx + i;
        const auto f =
            rotate_vector(tw[positive_modulo(ti, count)], -tf_delta);
        factors.insert(factors.end(), f.begin(), f.end());
      }
    }
    const int messages = int(factors.size() / slots);
    const Text prepared =
        e.prepare(e.encode_messages(factors, messages, level));
    fhe_context::StagePlaintexts sp;
    sp.count = messages;
    sp.level = level;
    const size_t bytes = size_t(e.words(prepared)) * 8;
    hip_check(hipMalloc(&sp.data, bytes), "diagonal allocation");
    hip_check(hipMemcpyAsync(sp.data, prepared.data, bytes,
                             hipMemcpyDeviceToDevice, ctx->stream),
              "copy");
    it = ctx->boot_diagonals.emplace(key, sp).first;
  }
  Text pts = e.wrap(static_cast<const int64_t *>(it->second.data),
                    it->second.count, level,
                    FHE_FLAG_NTT_FORM | FHE_FLAG_RESCALED);
  pts.count = 1;
  // each baby-step text prepared once for all of its products
  std::vector<Text> prepared;
  for (const Text &r : rotated)
    prepared.push_back(e.prepare(r));
  int next_pt = 0;
  std::vector<Text> sums(outer);
  std::vector<int> giant(outer);
  for (int o = 0; o < outer; o++) {
    const int nidx = o - outer / 2;
    giant[o] = e.normalized(inner * nidx * rotate_count, true);
    bool have_inner = false;
    for (int i = (o == 0 ? 1 : 0); i < inner; i++) {
      Text m = e.multiply_prepared(prepared[i], pts, next_pt++);
      if (!have_inner) {
        sums[o] = m;
        have_inner = true;
      } else {
        e.add_inplace(sums[o], m);
      }
    }
    require(have_inner, "Bootstrap: an empty butterfly row");
  }
  // giant steps: rotate(inner_sum, RotationKey, delta) of every row, each
  // the chain of fixed rotations of its delta's bits in the key's order
  // (src/ckks/engine_rotate.cpp:111-139).  The chains advance together key
  // by key, the rows that need a key rotated in one batched call: each
  // row's sequence of operations is unchanged.
  for (int o = 0; o < outer; o++)
    sums[o] = giant[o] == 0 ? e.intt(sums[o]) : e.key_switch_input(sums[o]);
  e.rotate_rows(sums, giant);
  Text outer_sum = sums[0];
  for (int o = 1; o < outer; o++)
    e.add_inplace(outer_sum, sums[o]);
  return outer_sum;
}

int log2_int(int64_t v) {
  int l = 0;
  while ((int64_t(1) << l) < v)
    l++;
  return l;
}

// coeff_to_slot with a RotationKey (src/ckks/engine_bootstrap.cpp:330-366)
const std::vector<Vectors> &cached_twiddles(fhe_context *ctx, int slots,
                                            int stages, bool to_slot) {
  const int log_slots = log2_int(slots);
  const auto key = std::make_tuple(log_slots, stages, to_slot);
  auto it = ctx->boot_twiddles.find(key);
  if (it == ctx->boot_twiddles.end())
    it = ctx->boot_twiddles
             .emplace(key, stage_twiddles(slots, log_slots, stages, to_slot))
             .first;
  return it->second;
}

std::pair<Text, Text> coeff_to_slot(TextEngine &e, const Text &x, int slots,
                                    int stages) {
  const int log_slots = log2_int(slots);
  const auto &tw = cached_twiddles(e.context(), slots, stages, true);
  int rotate_count = slots;
  Text current = x;
  for (int st = 0; st < stages; st++) {
    rotate_count >>= cts_log_radix(st, stages, log_slots);
    current = apply_butterfly(e, current, tw[st], rotate_count, true,
                              stages, st);
  }
  const Text conjugated = e.conjugate(current);
  Text real = e.add(current, conjugated, false);
  const Text subtracted = e.add(current, conjugated, true);
  Text imag = e.multiply_imaginary(subtracted, -1);
  return {real, imag};
}

// slot_to_coeff with a RotationKey (src/ckks/engine_bootstrap.cpp:405-434)
Text slot_to_coeff(TextEngine &e, const Text &real, const Text &imag,
                   int slots, int stages) {
  const int log_slots = log2_int(slots);
  const Text multiplied = e.multiply_imaginary(imag, 1);
  Text current = e.add(real, multiplied, false);
  const auto &tw = cached_twiddles(e.context(), slots, stages, false);
  int rotate_count = 1;
  for (int st = 0; st < stages; st++) {
    current = apply_butterfly(e, current, tw[st], rotate_count, false,
                              stages, st);
    rotate_count <<= stc_log_radix(st, stages, log_slots);
  }
  return current;
}

void require_bootstrap_keys(const fhe_bootstrap_keys *k) {
  require(k && k->relin && k->conjugation && k->rotation_count > 0 &&
              k->rotation && k->rotation_moduli_ids && k->rotation_deltas,
          "Bootstrap: keys missing");
  for (int i = 0; i < k->rotation_count; i++)
    require(k->rotation[i] != nullptr, "Bootstrap: rotation key missing");
}

} // namespace

namespace fhe_host {

Text round_text(TextEngine &e, const Text &x) {
  // round = x * ROUND_BOUND - modulo(x), modulo = double angle of sine
  // (src/ckks/engine_bootstrap.cpp:661-691)
  const Text decimal = double_angle(e, sine(e, x));
  const Text multiplied = e.multiply_int(x, kRoundBound);
  return e.add(multiplied, decimal, true);
}

void copy_out(fhe_context *ctx, TextEngine &e, const Text &t, int64_t *out,
              int *out_level, int *out_flags) {
  hip_check(hipMemcpyAsync(out, t.data, size_t(e.words(t)) * 8,
                           hipMemcpyDeviceToDevice, ctx->stream),
            "copy");
  *out_level = t.level;
  *out_flags = t.flags;
}

// The form of a bootstrap input: known data-flag bits only and a level on
// the chain (checked before any device work; a level off the chain would
// index the device tables out of range).
void require_boot_form(const fhe::CkksParams &P, int level, int flags,
                       int min_level, const char *op) {
  require((flags & ~(FHE_FLAG_NTT_FORM | FHE_FLAG_RESCALED)) == 0,
          std::string(op) + ": unknown data flags");
  require(level >= min_level && level <= P.full_level,
          std::string(op) + ": the level is out of range.");
}

// CKKSEngine::bootstrap (src/ckks/engine_bootstrap.cpp:712-754 with a
// RotationKey, 794-834 with a BootstrapKey) on `count` ciphertexts: level
// down to 0, mod_raise, the partial sums below full slots, coeff_to_slot,
// round of both parts in one batched EvalMod (each ciphertext follows its
// own sequence of operations), slot_to_coeff, the subtraction from the
// raised text and the normalizing integer product.  The steps that depend
// on the kind of rotation key come from `steps`.
Text bootstrap_texts(TextEngine &e, const Text &x,
                     const BootstrapSteps &steps) {
  fhe_context *ctx = e.context();
  const auto &P = ctx->params;
  const int slots = int(P.slot_count());
  const Text leveled = e.level_down(x, 0);
  Text raised = e.mod_raise(leveled);
  if (slots != ctx->n() / 2)
    steps.partial_sum(raised);
  const auto [real, imag] = steps.coeff_to_slot(raised);
  Text both = e.alloc(real.polys, real.level, real.flags, 2 * real.count);
  const int64_t part = e.words(real);
  hip_check(hipMemcpyAsync(both.data, real.data, size_t(part) * 8,
                           hipMemcpyDeviceToDevice, ctx->stream),
            "copy");
  hip_check(hipMemcpyAsync(both.data + part, imag.data, size_t(part) * 8,
                           hipMemcpyDeviceToDevice, ctx->stream),
            "copy");
  const Text rounded = round_text(e, both);
  Text rounded_real = rounded, rounded_imag = rounded;
  rounded_real.count = rounded_imag.count = real.count;
  rounded_imag.data = rounded.data + e.words(rounded_real);
  const Text coeff = steps.slot_to_coeff(rounded_real, rounded_imag);
  const Text bootstrapped = e.add(raised, coeff, true);
  // CKKSCore::get_normalize_ratio (src/ckks/core.cpp:1121-1125)
  const int64_t ratio = static_cast<int64_t>(
      std::round(double(P.q[P.chain_count - 1]) / P.scale));
  return e.multiply_int(bootstrapped, ratio);
}

} // namespace fhe_host

extern "C" {

int fhe_bootstrap_twiddles(int log_slots, int stage_count, int to_slot,
                           int stage, double *out, int64_t capacity,
                           int64_t *count) {
  return guarded([&] {
    require(out && count && log_slots >= 1 && log_slots <= 16 &&
                stage_count >= 1 && stage >= 0 && stage < stage_count,
            "Bootstrap twiddles: arguments out of range");
    const int slots = 1 << log_slots;
    const auto tw = stage_twiddles(slots, log_slots, stage_count, to_slot != 0);
    const auto &st = tw[stage];
    *count = int64_t(st.size());
    require(capacity >= int64_t(st.size()) * slots * 2,
            "Bootstrap twiddles: capacity too small");
    for (size_t r = 0; r < st.size(); r++)
      std::memcpy(out + r * slots * 2, st[r].data(), size_t(slots) * 16);
  });
}

int fhe_bootstrap_key_stage_layout(const fhe_params *params, int stage_count,
                                   int inner_loop_shift, int to_slot,
                                   int stage, int *level, int *inner_count,
                                   int *outer_count, int *inner_deltas,
                                   int *outer_deltas, int capacity) {
  return guarded([&] {
    require(level && inner_count && outer_count, "null argument");
    const auto P = resolve_params(params, false);
    int log_slots = 0;
    while ((int64_t(1) << log_slots) < P.slot_count())
      log_slots++;
    require(bootstrap_stage_count_valid(log_slots, inner_loop_shift,
                                        stage_count),
            "Bootstrap key layout: invalid stage count.");
    require(stage >= 0 && stage < stage_count,
            "Bootstrap key layout: stage out of range");
    const StageLayout s = stage_layout(P, stage_count, inner_loop_shift,
                                       to_slot != 0, stage);
    require(s.level >= 1, "Bootstrap key layout: too few levels");
    *level = s.level;
    *inner_count = s.inner;
    *outer_count = s.outer;
    if (inner_deltas || outer_deltas)
      require(capacity >= std::max(s.inner, s.outer),
              "Bootstrap key layout: capacity too small");
    if (inner_deltas)
      std::copy(s.inner_deltas.begin(), s.inner_deltas.end(), inner_deltas);
    if (outer_deltas)
      std::copy(s.outer_deltas.begin(), s.outer_deltas.end(), outer_deltas);
  });
}

int fhe_bootstrap_key_rotation_layout(const fhe_params *params,
                                      int stage_count, int inner_loop_shift,
                                      int *deltas, int *moduli_ids,
                                      int capacity, int *count) {
  return guarded([&] {
    require(count, "null argument");
    const auto P = resolve_params(params, false);
    const int slots = int(P.slot_count());
    int log_slots = 0;
    while ((1 << log_slots) < slots)
      log_slots++;
    require(bootstrap_stage_count_valid(log_slots, inner_loop_shift,
                                        stage_count),
            "Bootstrap key layout: invalid stage count.");
    // MAKE_ROTATION_CONTEXT's order: partial sums at the maximum
    // key-switching level, then every stage's inner and outer deltas at the
    // stage's level; a delta keeps the level of its first appearance
    std::vector<std::pair<int, int>> out;
    std::set<int> seen;
    auto add = [&](int delta, int level) {
      if (delta != 0 && seen.insert(delta).second)
        out.push_back({delta, P.moduli_id_of_level(level)});
    };
    const int top = P.max_key_switching_level();
    for (int d = slots; d < P.n / 2; d <<= 1)
      add(d, top);
    for (int to_slot = 1; to_slot >= 0; to_slot--)
      for (int st = 0; st < stage_count; st++) {
        const StageLayout s =
            stage_layout(P, stage_count, inner_loop_shift, to_slot != 0, st);
        for (int d : s.inner_deltas)
          add(d, s.level);
        for (int d : s.outer_deltas)
          add(d, s.level);
      }
    *count = int(out.size());
    if (deltas || moduli_ids) {
      require(capacity >= int(out.size()),
              "Bootstrap key layout: capacity too small");
      for (size_t i = 0; i < out.size(); i++) {
        if (deltas)
          deltas[i] = out[i].first;
        if (moduli_ids)
          moduli_ids[i] = out[i].second;
      }
    }
  });
}

int fhe_bootstrap_key_encode_stage(fhe_context *ctx, int stage_count,
                                   int inner_loop_shift, int to_slot,
                                   int stage, int64_t *out,
                                   int64_t capacity_words, int *count) {
  return guarded([&] {
    require(ctx && count, "null argument");
    const auto &P = ctx->params;
    const int slots = int(P.slot_count());
    const int log_slots = log2_int(slots);
    require(bootstrap_stage_count_valid(log_slots, inner_loop_shift,
                                        stage_count),
            "Bootstrap key stage: invalid stage count.");
    require(stage >= 0 && stage < stage_count,
            "Bootstrap key stage: stage out of range");
    // make_bootstrap_stage (src/ckks/engine_bootstrap.cpp:8-69)
    const StageLayout s = stage_layout(P, stage_count, inner_loop_shift,
                                       to_slot != 0, stage);
    require(s.level >= 1, "Bootstrap key stage: too few levels");
    const int m = P.moduli_id_of_level(s.level) + 1; // after the rescale
    const int ext = P.prime_count - m;
    const int messages = s.outer * s.inner - 1;
    *count = messages;
    if (!out)
      return;
    const int64_t N = ctx->n();
    require(capacity_words >= int64_t(messages) * ext * N,
            "Bootstrap key stage: capacity too small");
    ctx->require_not_capturing("Bootstrap key stage");
    const auto &tw = cached_twiddles(ctx, slots, stage_count, to_slot != 0);
    const Vectors &st = tw[stage];
    const int twiddle_count = int(st.size());
    require((twiddle_count + 1) / s.inner == s.outer,
            "Bootstrap key stage: loop counts differ from the twiddles");
    std::vector<Complex> factors;
    factors.reserve(size_t(messages) * slots);
    for (int o = 0; o < s.outer; o++) {
      const int nidx = o - s.outer / 2;
      for (int i = (o == 0 ? 1 : 0); i < s.inner; i++) {
        const int ti = s.inner * nidx + i;
        const auto f = rotate_vector(st[positive_modulo(ti, twiddle_count)],
                                     -s.outer_deltas[o]);
        factors.insert(factors.end(), f.begin(), f.end());
      }
    }
    // encode at the stage's level with the special primes, rescale, NTT
    TextEngine e(ctx, nullptr, 0);
    const size_t bytes = factors.size() * sizeof(Complex);
    const auto dv = e.take(bytes);
    hip_check(hipMemcpyAsync(dv.get(), factors.data(), bytes,
                             hipMemcpyHostToDevice, ctx->stream),
              "message copy");
    const auto enc = e.take(size_t(messages) * (ext + 1) * N * 8);
    int64_t *encoded = static_cast<int64_t *>(enc.get());
    encode_texts(ctx, static_cast<const double *>(dv.get()), true, messages,
                 s.level, true, encoded);
    ctx->stage("bootstrap_key_stage", [&] {
      fhe::launch_rescale(ctx->tables, encoded, out, messages, ext + 1, m - 1,
                          ctx->stream);
      fhe::ntt_forward(ctx->tables, out, out, int64_t(messages) * ext, ext, m,
                       ctx->tables.r2, ctx->stream);
    });
    ctx->check_launch("Bootstrap key stage");
  });
}

int fhe_coeff_to_slot(fhe_context *ctx, int count, const int64_t *in,
                      int level, int flags, const fhe_bootstrap_keys *keys,
                      int stage_count, int64_t *real, int64_t *imag,
                      int *out_level, int *out_flags) {
  return guarded([&] {
    require(ctx && in && real && imag && out_level && out_flags,
            "null argument");
    require_bootstrap_keys(keys);
    require(stage_count >= 1, "Coeff to slot: stage count");
    require(count > 0, "Bootstrap: empty batch");
    require_boot_form(ctx->params, level, flags, 1, "Coeff to slot");
    TextEngine e(ctx, keys->relin, keys->relin_moduli_id);
    e.keys = keys;
    e.count = count;
    const Text x = e.wrap(in, 2, level, flags);
    const auto [r, i] =
        coeff_to_slot(e, x, int(ctx->params.slot_count()), stage_count);
    int l2, f2;
    copy_out(ctx, e, r, real, out_level, out_flags);
    copy_out(ctx, e, i, imag, &l2, &f2);
    require(l2 == *out_level && f2 == *out_flags, "Coeff to slot: forms");
    ctx->check_launch("Coeff to slot");
  });
}

int fhe_slot_to_coeff(fhe_context *ctx, int count, const int64_t *real,
                      const int64_t *imag, int level, int flags,
                      const fhe_bootstrap_keys *keys, int stage_count,
                      int64_t *out, int *out_level, int *out_flags) {
  return guarded([&] {
    require(ctx && real && imag && out && out_level && out_flags,
            "null argument");
    require_bootstrap_keys(keys);
    require(stage_count >= 1, "Slot to coeff: stage count");
    require(count > 0, "Bootstrap: empty batch");
    require_boot_form(ctx->params, level, flags, 1, "Slot to coeff");
    TextEngine e(ctx, keys->relin, keys->relin_moduli_id);
    e.keys = keys;
    e.count = count;
    const Text r = e.wrap(real, 2, level, flags);
    const Text i = e.wrap(imag, 2, level, flags);
    const Text c =
        slot_to_coeff(e, r, i, int(ctx->params.slot_count()), stage_count);
    copy_out(ctx, e, c, out, out_level, out_flags);
    ctx->check_launch("Slot to coeff");
  });
}

int fhe_bootstrap(fhe_context *ctx, int count, const int64_t *in, int level,
                  int flags,
                  const fhe_bootstrap_keys *keys, int stage_count,
                  int64_t *out, int *out_level, int *out_flags) {
  return guarded([&] {
    require(ctx && in && out && out_level && out_flags, "null argument");
    require_bootstrap_keys(keys);
    require(keys->to_sparse && keys->to_dense, "Bootstrap: keys missing");
    require(!(flags & FHE_FLAG_NTT_FORM) && stage_count >= 1,
            "Bootstrap: an INTT-form duo and a positive stage count");
    const auto &P = ctx->params;
    require(count > 0, "Bootstrap: empty batch");
    require_boot_form(P, level, flags, 0, "Bootstrap");
    TextEngine e(ctx, keys->relin, keys->relin_moduli_id);
    e.keys = keys;
    e.count = count;
    const int slots = int(P.slot_count());
    BootstrapSteps steps;
    // partial_sum_inplace (src/ckks/engine_bootstrap.cpp:315-329)
    steps.partial_sum = [&](Text &raised) {
      for (int delta = slots; delta < ctx->n() / 2; delta <<= 1) {
        const Text r = e.rotate(raised, delta, false);
        e.add_inplace(raised, r);
      }
    };
    steps.coeff_to_slot = [&](const Text &raised) {
      return coeff_to_slot(e, raised, slots, stage_count);
    };
    steps.slot_to_coeff = [&](const Text &real, const Text &imag) {
      return slot_to_coeff(e, real, imag, slots, stage_count);
    };
    const Text result = bootstrap_texts(e, e.wrap(in, 2, level, flags), steps);
    copy_out(ctx, e, result, out, out_level, out_flags);
    ctx->check_launch("Bootstrap");
  });
}

int fhe_round(fhe_context *ctx, int count, const int64_t *in, int level,
              const int64_t *key, int key_moduli_id, int64_t *out,
              int *out_level, int *out_flags) {
  return guarded([&] {
    require(ctx && in && key && out && out_level && out_flags,
            "null argument");
    require(level >= 1 && level <= ctx->params.full_level,
            "Round: the level must be positive.");
    require(count > 0, "Round: empty batch");
    TextEngine e(ctx, key, key_moduli_id);
    e.count = count;
    const Text x = e.wrap(in, 2, level, 0);
    copy_out(ctx, e, round_text(e, x), out, out_level, out_flags);
    ctx->check_launch("Round");
  });
}

} // extern "C"
