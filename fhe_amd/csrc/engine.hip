// Engine-level compositions on the device (CKKSEngine's form bookkeeping,
// EvalMod, the slot transforms and the bootstrap), host C++ over the device
// operations of the C ABI.
#include "context.h"

namespace {

// A ciphertext on the device with the metadata of the reference's
// UnitCiphertext that the limbs depend on (the scale does not enter them).
struct Text {
  int64_t *data = nullptr;
  int polys = 2;
  int level = 0;
  int flags = 0;
  int count = 1;              // ciphertexts (or plaintexts) side by side
  std::shared_ptr<void> hold; // returns the buffer to the context's pool
};


// CKKSEngine's operations on device texts, each with the engine's form
// handling (src/ckks/engine_*.cpp), on `count` ciphertexts at once (the
// reference runs the same operations one ciphertext at a time).  Results
// are new device buffers that go back to the context's pool when their last
// Text is gone: later work on the one stream reuses them in order.
class TextEngine {
public:
  TextEngine(fhe_context *c, const int64_t *rlk, int rlk_mid)
      : ctx(c), key(rlk), key_mid(rlk_mid) {}
  ~TextEngine() { (void)hipStreamSynchronize(ctx->stream); }

  int count = 1;

  fhe_context *context() const { return ctx; }
  int limbs(int level, int flags) const {
    return ctx->params.limbs_at(
        ctx->moduli_id_of(level, (flags & FHE_FLAG_RESCALED) != 0));
  }
  int64_t words(const Text &t) const {
    return int64_t(t.count) * t.polys * limbs(t.level, t.flags) * ctx->n();
  }
  std::shared_ptr<void> take(size_t bytes) {
    fhe_context *c = ctx;
    return std::shared_ptr<void>(c->pool_take(bytes), [c, bytes](void *p) {
      c->text_pool.emplace(bytes, p);
    });
  }
  Text alloc(int polys, int level, int flags, int n_texts = 0) {
    Text t;
    t.polys = polys;
    t.level = level;
    t.flags = flags;
    t.count = n_texts > 0 ? n_texts : count;
    t.hold = take(size_t(words(t)) * 8);
    t.data = static_cast<int64_t *>(t.hold.get());
    return t;
  }
  Text wrap(const int64_t *data, int polys, int level, int flags) {
    Text t;
    t.data = const_cast<int64_t *>(data);
    t.polys = polys;
    t.level = level;
    t.flags = flags;
    t.count = count;
    return t;
  }
  Text clone(const Text &a) {
    Text t = alloc(a.polys, a.level, a.flags, a.count);
    hip_check(hipMemcpyAsync(t.data, a.data, size_t(words(a)) * 8,
                             hipMemcpyDeviceToDevice, ctx->stream),
              "copy");
    return t;
  }
  // intt_ (src/ckks/engine_intt.cpp): a copy in the coefficient domain
  Text intt(const Text &a) {
    if (!(a.flags & FHE_FLAG_NTT_FORM))
      return a;
    Text t = clone(a);
    t.flags &= ~FHE_FLAG_NTT_FORM;
    const int l = limbs(a.level, a.flags);
    ok(fhe_intt(ctx, t.data, a.count * a.polys, l,
                ctx->moduli_id_of(a.level, a.flags & FHE_FLAG_RESCALED),
                1));
    return t;
  }
  // multiply(ct, ct[, rlk]), square (src/ckks/engine_multiply.cpp:39-199)
  Text multiply(const Text &a, const Text &b, bool relin) {
    require(a.polys == 2 && b.polys == 2, "Multiply: duos expected");
    const int level = std::min(a.level, b.level);
    Text out = alloc(relin ? 2 : 3, level - 1, relin ? 0 : FHE_FLAG_NTT_FORM,
                     a.count);
    ok(fhe_multiply(ctx, a.count, a.data, a.level, a.flags, b.data, b.level,
                    b.flags, relin ? key : nullptr, key_mid, out.data));
    return out;
  }
  // add / subtract (src/ckks/engine_add.cpp:82-110,
  // src/ckks/engine_subtract.cpp:5-35), forms as match_form leaves them
  Text add(const Text &a, const Text &b, bool subtract) {
    const int level = std::min(a.level, b.level);
    const bool ra = (a.flags & FHE_FLAG_RESCALED) && a.level == level;
    const bool rb = (b.flags & FHE_FLAG_RESCALED) && b.level == level;
    const int flags = (a.flags & FHE_FLAG_NTT_FORM) |
                      (ra || rb ? FHE_FLAG_RESCALED : 0);
    Text out = alloc(std::max(a.polys, b.polys), level, flags, a.count);
    ok(fhe_add_texts(ctx, a.count, a.data, a.polys, a.level, a.flags, b.data,
                     b.polys, b.level, b.flags, subtract ? 1 : 0, out.data));
    return out;
  }
  // add_inplace of one form (CKKSCore::add_inplace,
  // src/ckks/core.cpp:498-512)
  void add_inplace(Text &x, const Text &y) {
    require(x.polys == y.polys && x.level == y.level && x.flags == y.flags,
            "Add inplace: the forms differ");
    const int l = limbs(x.level, x.flags);
    fhe::launch_elementwise(ctx->tables, fhe::kEwAddReduce, x.data, y.data,
                            x.data, int64_t(x.count) * x.polys * l, l,
                            ctx->moduli_id_of(x.level,
                                              x.flags & FHE_FLAG_RESCALED),
                            nullptr, ctx->stream);
  }
  // add(ct, double) of a coefficient-domain, not rescaled text
  // (src/ckks/engine_add.cpp:518-536 with CKKSCore::add(ct, double))
  Text add_double(const Text &a, double value) {
    require(!(a.flags & (FHE_FLAG_NTT_FORM | FHE_FLAG_RESCALED)),
            "Add: a scalar joins only an INTT-form, not rescaled text here");
    Text out = alloc(a.polys, a.level, a.flags, a.count);
    ok(fhe_add_double(ctx, a.data, out.data, a.count, a.polys, a.level,
                      value));
    return out;
  }
  // multiply(ct, double) (src/ckks/engine_multiply.cpp:418-432): intt_,
  // then the scalar product and rescale to level - 1, not rescaled
  Text multiply_double(const Text &a, double value) {
    require(!(a.flags & FHE_FLAG_RESCALED) && a.level >= 1,
            "Multiply: the text should not be rescaled and have a positive "
            "level.");
    const Text in = intt(a);
    Text out = alloc(a.polys, a.level - 1, in.flags, a.count);
    ok(fhe_multiply_double(ctx, in.data, out.data, a.count * a.polys, a.level,
                           value));
    return out;
  }
  // multiply(ct, int64_t) (src/ckks/engine_multiply.cpp:443-457)
  Text multiply_int(const Text &a, int64_t value) {
    require(!(a.flags & FHE_FLAG_RESCALED),
            "Multiply: the text should not be rescaled.");
    const Text in = intt(a);
    Text out = alloc(a.polys, a.level, in.flags, a.count);
    ok(fhe_multiply_int(ctx, in.data, out.data, a.count * a.polys,
                        ctx->moduli_id_of(a.level, false), value));
    return out;
  }
  // level_down (src/ckks/engine_level_down.cpp:5-37)
  Text level_down(const Text &a, int level) {
    if (a.level == level)
      return clone(a);
    Text out = alloc(a.polys, level, a.flags & ~FHE_FLAG_RESCALED, a.count);
    const int64_t t_l = words(a);
    int64_t *t0 = ctx->scratch(size_t(2 * t_l) * 8);
    ctx->transform_into(a.data, int64_t(a.count) * a.polys, a.level, a.flags,
                        false, level,
                        false, (a.flags & FHE_FLAG_NTT_FORM) != 0, out.data,
                        t0, t0 + t_l);
    return out;
  }
  // relinearize (src/ckks/engine_relinearize.cpp:13-28): ntt_, then
  // CKKSCore::relinearize
  Text relinearize(const Text &trio) {
    require(trio.polys == 3, "Relinearize: a trio expected");
    const bool rescaled = (trio.flags & FHE_FLAG_RESCALED) != 0;
    const int mid = ctx->moduli_id_of(trio.level, rescaled);
    Text in = trio;
    if (!(trio.flags & FHE_FLAG_NTT_FORM)) {
      in = clone(trio);
      in.flags |= FHE_FLAG_NTT_FORM;
      ok(fhe_ntt(ctx, in.data, 3 * trio.count, limbs(trio.level, trio.flags),
                 mid, 1));
    }
    ctx->check_key(mid, key_mid, "Relinearize");
    Text out = alloc(2, trio.level, trio.flags & ~FHE_FLAG_NTT_FORM,
                     trio.count);
    ctx->relinearize_into(trio.count, mid, in.data, key, key_mid, out.data);
    return out;
  }
  // weighted_sum (src/ckks/engine_advanced.cpp:228-269)
  Text weighted_sum(const std::vector<Text> &basis,
                    const std::vector<double> &weights) {
    const int degree = int(weights.size()) - 1;
    Text result;
    bool have = false;
    for (int index = degree; index > 0; index--) {
      const double weight = weights[index];
      if (weight == 0)
        continue;
      Text product = multiply_double(basis[index - 1], weight);
      if (!have) {
        result = product;
        have = true;
        continue;
      }
      result = add(result, product, false);
    }
    require(have, "Weighted sum: every weight is zero");
    if (weights[0] != 0)
      result = add_double(result, weights[0]);
    return result;
  }

  // --- operations with the bootstrap keys ---------------------------------
  const fhe_bootstrap_keys *keys = nullptr;

  bool at_top(const Text &a) const {
    return a.level == ctx->params.full_level &&
           !(a.flags & FHE_FLAG_RESCALED);
  }
  // the one-more-level chain's top text goes one level down before a key
  // switch (src/ckks/engine_rotate.cpp:63-67), then leaves the NTT domain
  Text key_switch_input(const Text &a) {
    return intt(at_top(a) ? level_down(a, a.level - 1) : a);
  }
  // core rotate of a coefficient-domain text by one fixed rotation key
  Text rotate_core(const Text &a, int index) {
    const int mid = ctx->moduli_id_of(a.level, a.flags & FHE_FLAG_RESCALED);
    Text out = alloc(a.polys, a.level, a.flags, a.count);
    ok(fhe_rotate(ctx, a.count, mid, a.data, keys->rotation[index],
                  keys->rotation_moduli_ids[index], keys->rotation_deltas[index],
                  out.data));
    return out;
  }
  int normalized(int delta, bool by_slots) const {
    const int m = by_slots ? int(ctx->params.slot_count())
                           : int(ctx->n() / 2);
    return ((delta % m) + m) % m;
  }
  // rotate(ct, RotationKey, delta) (src/ckks/engine_rotate.cpp:111-139):
  // one fixed rotation per set bit, in the key's order
  Text rotate(const Text &a, int delta, bool by_slots = true) {
    const int d = normalized(delta, by_slots);
    if (d == 0)
      return a.flags & FHE_FLAG_NTT_FORM ? intt(a) : clone(a);
    Text current = a;
    bool first = true;
    for (int k = 0; k < keys->rotation_count; k++) {
      if ((d & keys->rotation_deltas[k]) == 0)
        continue;
      current = rotate_core(first ? key_switch_input(current) : current, k);
      first = false;
    }
    return current;
  }
  // Advances every row's chain of fixed rotations (the set bits of its
  // normalized delta, in the key's order) key by key, the rows that use a
  // key batched into one call; rows are coefficient-domain texts of one
  // form, replaced by their rotations.
  void rotate_rows(std::vector<Text> &rows, const std::vector<int> &deltas) {
    const int64_t row_words = words(rows[0]);
    for (int k = 0; k < keys->rotation_count; k++) {
      std::vector<int> members;
      for (size_t r = 0; r < rows.size(); r++)
        if (deltas[r] & keys->rotation_deltas[k])
          members.push_back(int(r));
      if (members.empty())
        continue;
      const Text &first = rows[members[0]];
      Text batch = alloc(first.polys, first.level, first.flags,
                         first.count * int(members.size()));
      for (size_t m = 0; m < members.size(); m++)
        hip_check(hipMemcpyAsync(batch.data + m * row_words,
                                 rows[members[m]].data, size_t(row_words) * 8,
                                 hipMemcpyDeviceToDevice, ctx->stream),
                  "copy");
      const Text rotated = rotate_core(batch, k);
      for (size_t m = 0; m < members.size(); m++) {
        Text view = rotated;
        view.count = first.count;
        view.data = rotated.data + m * row_words;
        rows[members[m]] = view;
      }
    }
  }
  // rotate_batch(ct, RotationKey, deltas) (src/ckks/engine_rotate.cpp:
  // 305-393): chains of fixed rotations from the prepared input, shared
  // prefixes computed once
  std::vector<Text> rotate_batch(const Text &a, const std::vector<int> &deltas) {
    const Text base = key_switch_input(a);
    std::map<int, Text> done{{0, base}};
    std::vector<int> targets;
    for (int delta : deltas)
      targets.push_back(normalized(delta, true));
    const int64_t row_words = words(base);
    // keys in the key's order: every rotation by key k extends a prefix
    // of lower bits, all of them done already; the rotations by one key
    // run as one batched call
    for (int k = 0; k < keys->rotation_count; k++) {
      const int kd = keys->rotation_deltas[k];
      std::vector<int> prevs;
      for (int d : targets) {
        if (!(d & kd))
          continue;
        const int prev = d & (kd - 1);
        if (!done.count(prev + kd) &&
            std::find(prevs.begin(), prevs.end(), prev) == prevs.end())
          prevs.push_back(prev);
      }
      if (prevs.empty())
        continue;
      Text batch = alloc(base.polys, base.level, base.flags,
                         base.count * int(prevs.size()));
      for (size_t m = 0; m < prevs.size(); m++)
        hip_check(hipMemcpyAsync(batch.data + m * row_words,
                                 done.at(prevs[m]).data,
                                 size_t(row_words) * 8,
                                 hipMemcpyDeviceToDevice, ctx->stream),
                  "copy");
      const Text rotated = rotate_core(batch, k);
      for (size_t m = 0; m < prevs.size(); m++) {
        Text view = rotated;
        view.count = base.count;
        view.data = rotated.data + m * row_words;
        done[prevs[m] + kd] = view;
      }
    }
    std::vector<Text> out;
    for (int d : targets)
      out.push_back(done.at(d));
    return out;
  }
  // encode_unit of `count` complex messages [count][slots] at `level`
  // (src/ckks/engine_encode.cpp:10-36), one batched device encode
  Text encode_messages(const std::vector<std::complex<double>> &m,
                       int messages, int level) {
    const size_t bytes = m.size() * sizeof(std::complex<double>);
    const auto dv = take(bytes);
    hip_check(hipMemcpyAsync(dv.get(), m.data(), bytes, hipMemcpyHostToDevice,
                             ctx->stream),
              "message copy");
    // one plaintext per message, shared by every ciphertext
    Text pt = alloc(messages, level, 0, 1);
    ok(fhe_encode(ctx, static_cast<const double *>(dv.get()), 1, messages,
                  level, pt.data));
    // the host messages may go once the stream has consumed them
    hip_check(hipStreamSynchronize(ctx->stream), "message sync");
    return pt;
  }
  // prepare_for_multiply (src/ckks/engine_multiply.cpp:13-27) of a text at
  // its own level: rescale_ and ntt_; every polynomial of a batch of
  // plaintexts in one pass
  Text prepare(const Text &a) {
    const bool rescaled = (a.flags & FHE_FLAG_RESCALED) != 0;
    const bool ntt = (a.flags & FHE_FLAG_NTT_FORM) != 0;
    if (rescaled && ntt)
      return a;
    Text out = alloc(a.polys, a.level, FHE_FLAG_NTT_FORM | FHE_FLAG_RESCALED,
                     a.count);
    const int64_t t_l = words(a);
    int64_t *t0 = ctx->scratch(size_t(2 * t_l) * 8);
    const int64_t *r = ctx->prepare_for_multiply_into(
        a.data, int64_t(a.count) * a.polys, a.level, a.flags, a.level,
        out.data, t0, t0 + t_l);
    require(r == out.data, "Prepare: unexpected form");
    return out;
  }
  // CKKSCore::multiply(ct, pt) (src/ckks/core.cpp:257-282) of a prepared
  // duo and plaintext `index` of a prepared batch: the product of
  // multiply(ct, message) (src/ckks/engine_multiply.cpp:292-305), NTT form,
  // not rescaled, one level down
  Text multiply_prepared(const Text &ct, const Text &pts, int index) {
    require(ct.level == pts.level && ct.flags == pts.flags &&
                ct.flags == (FHE_FLAG_NTT_FORM | FHE_FLAG_RESCALED),
            "Multiply: prepared operands expected");
    const int mid = ctx->moduli_id_of(ct.level, true);
    const int l = ctx->params.limbs_at(mid);
    const int64_t poly = int64_t(l) * ctx->n();
    Text out = alloc(ct.polys, ct.level - 1, FHE_FLAG_NTT_FORM, ct.count);
    ctx->stage("mont_mult", [&] {
      fhe::launch_mont_mult_broadcast(
          ctx->tables, ct.data, pts.data + index * poly, out.data,
          int64_t(ct.count) * ct.polys * l, l, mid, ctx->stream);
    });
    return out;
  }
  // conjugate (src/ckks/engine_conjugate.cpp)
  Text conjugate(const Text &a) {
    const Text in = key_switch_input(a);
    const int mid = ctx->moduli_id_of(in.level, in.flags & FHE_FLAG_RESCALED);
    Text out = alloc(in.polys, in.level, in.flags, in.count);
    ok(fhe_conjugate(ctx, in.count, mid, in.data, keys->conjugation,
                     keys->conjugation_moduli_id, out.data));
    return out;
  }
  // multiply_imaginary of a coefficient-domain, not rescaled text
  // (src/ckks/engine_multiply.cpp:468-491)
  Text multiply_imaginary(const Text &a, int64_t value) {
    require(!(a.flags & (FHE_FLAG_NTT_FORM | FHE_FLAG_RESCALED)),
            "Multiply imaginary: an INTT-form, not rescaled text here");
    Text out = alloc(a.polys, a.level, a.flags, a.count);
    ok(fhe_multiply_imaginary(ctx, a.count, ctx->moduli_id_of(a.level, false),
                              a.data, value, out.data));
    return out;
  }
  // mod_raise_ (src/ckks/engine_mod_raise.cpp:5-34) of a level-0 text
  Text mod_raise(const Text &a) {
    require(a.level == 0 && a.flags == 0,
            "Mod raise: an INTT-form, not rescaled level-0 text");
    const auto &P = ctx->params;
    const int level = P.max_key_switching_level();
    Text out = alloc(2, level, 0, a.count);
    ok(fhe_mod_raise(ctx, a.count, a.data, keys->to_sparse,
                     keys->to_sparse_moduli_id, keys->to_dense,
                     keys->to_dense_moduli_id, out.data));
    return out;
  }

private:
  fhe_context *ctx;
  const int64_t *key;
  int key_mid;
};

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
  std::vector<Text> baby(kBabyDegree);
  baby[0] = e.clone(x);
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
  giant[0] = e.clone(baby.back());
  for (int gi = 1; gi < kGiantDegree; gi++) {
    Text m = e.multiply(giant[gi - 1], giant[gi - 1], true);
    e.add_inplace(m, m);
    giant[gi] = e.add_double(m, -1);
  }
  Text ps = e.clone(giant[0]);
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
  Text current = e.clone(x);
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

void require_bootstrap_keys(const fhe_bootstrap_keys *k) {
  require(k && k->relin && k->conjugation && k->rotation_count > 0 &&
              k->rotation && k->rotation_moduli_ids && k->rotation_deltas,
          "Bootstrap: keys missing");
  for (int i = 0; i < k->rotation_count; i++)
    require(k->rotation[i] != nullptr, "Bootstrap: rotation key missing");
}

} // namespace

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
    // CKKSEngine::bootstrap with a RotationKey
    // (src/ckks/engine_bootstrap.cpp:712-754)
    const Text x = e.wrap(in, 2, level, flags);
    const Text leveled = e.level_down(x, 0);
    Text raised = e.mod_raise(leveled);
    if (slots != ctx->n() / 2) {
      // partial_sum_inplace (src/ckks/engine_bootstrap.cpp:315-329)
      for (int delta = slots; delta < ctx->n() / 2; delta <<= 1) {
        const Text r = e.rotate(raised, delta, false);
        e.add_inplace(raised, r);
      }
    }
    const auto [real, imag] = coeff_to_slot(e, raised, slots, stage_count);
    // the two parts through one batched EvalMod: every ciphertext of the
    // batch follows its own sequence of operations
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
    const Text coeff =
        slot_to_coeff(e, rounded_real, rounded_imag, slots, stage_count);
    const Text bootstrapped = e.add(raised, coeff, true);
    // CKKSCore::get_normalize_ratio (src/ckks/core.cpp:1121-1125)
    const int64_t ratio = static_cast<int64_t>(
        std::round(double(P.q[P.chain_count - 1]) / P.scale));
    const Text result = e.multiply_int(bootstrapped, ratio);
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
