// CKKSEngine's form bookkeeping over device texts: engine-level operations
// composed of the C ABI's device operations, shared by the bootstrap, the
// advanced compositions and the serving loop.
#pragma once

#include <functional>

#include "context.h"

namespace fhe_host {

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
      : ctx(c), key(rlk), key_mid(rlk_mid), sync_on_exit(!c->capturing) {}
  ~TextEngine() {
    if (sync_on_exit)
      (void)hipStreamSynchronize(ctx->stream);
  }

  int count = 1;
  // the destructor waits for the stream (not while the stream is captured
  // into a graph; callers that poll events instead turn it off; pooled
  // buffers are reused in stream order either way)
  bool sync_on_exit;

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
  // rescale_ (src/ckks/engine_rescale.cpp:5-31) of a coefficient-domain
  // text: the text itself when it is rescaled already
  Text rescale(const Text &a) {
    if (a.flags & FHE_FLAG_RESCALED)
      return a;
    require(!(a.flags & FHE_FLAG_NTT_FORM),
            "Rescale: a coefficient-domain text here");
    Text out = alloc(a.polys, a.level, FHE_FLAG_RESCALED, a.count);
    ok(fhe_rescale(ctx, a.data, out.data, a.count * a.polys,
                   ctx->moduli_id_of(a.level, false)));
    return out;
  }
  // intt_ (src/ckks/engine_intt.cpp): a copy in the coefficient domain,
  // transformed out of place
  Text intt(const Text &a) {
    if (!(a.flags & FHE_FLAG_NTT_FORM))
      return a;
    Text t = alloc(a.polys, a.level, a.flags & ~FHE_FLAG_NTT_FORM, a.count);
    const int l = limbs(a.level, a.flags);
    fhe::ntt_inverse(ctx->tables, a.data, t.data,
                     int64_t(a.count) * a.polys * l, l,
                     ctx->moduli_id_of(a.level, a.flags & FHE_FLAG_RESCALED),
                     fhe::kInttExit, ctx->stream);
    ctx->check_launch("INTT");
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
      // to the NTT form out of place
      in = alloc(3, trio.level, trio.flags | FHE_FLAG_NTT_FORM, trio.count);
      const int l = limbs(trio.level, trio.flags);
      fhe::ntt_forward(ctx->tables, trio.data, in.data,
                       int64_t(3) * trio.count * l, l, mid, ctx->tables.r2,
                       ctx->stream);
      ctx->check_launch("NTT");
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
    // one pass over all terms when they share a level and the
    // coefficient-domain, not rescaled form: the same limbs as the calls
    {
      std::vector<int> used;
      for (int index = degree; index > 0; index--)
        if (weights[index] != 0)
          used.push_back(index - 1);
      const Text *first = used.empty() ? nullptr : &basis[used[0]];
      bool fused = first && int(used.size()) <= fhe::kWsMaxTerms &&
                   first->flags == 0 && first->level >= 1 &&
                   limbs(first->level, 0) <= fhe::kWsMaxLimbs;
      for (int b : used)
        fused = fused && basis[b].level == first->level &&
                basis[b].flags == first->flags &&
                basis[b].polys == first->polys &&
                basis[b].count == first->count;
      if (fused) {
        const int level = first->level;
        std::vector<const int64_t *> inputs;
        std::vector<std::vector<int64_t>> constants;
        for (int index = degree; index > 0; index--) {
          if (weights[index] == 0)
            continue;
          inputs.push_back(basis[index - 1].data);
          constants.push_back(
              ctx->params.mult_multiplier(weights[index], level));
        }
        std::vector<const int64_t *> cptr;
        for (const auto &c : constants) {
          require(int(c.size()) == limbs(level, 0),
                  "Weighted sum: multiplier size");
          cptr.push_back(c.data());
        }
        Text result = alloc(first->polys, level - 1, 0, first->count);
        ctx->stage("weighted_sum", [&] {
          fhe::launch_weighted_sum(
              ctx->tables, inputs.data(), cptr.data(), int(inputs.size()),
              result.data, first->count * first->polys, limbs(level, 0),
              ctx->moduli_id_of(level, false), ctx->stream);
        });
        if (weights[0] != 0)
          result = add_double(result, weights[0]);
        return result;
      }
    }
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
    ctx->require_not_capturing("Encode");
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

// The engine's advanced compositions (advanced.hip).
std::vector<Text> make_power_basis(TextEngine &e, const Text &x,
                                   int max_power);
Text evaluate_polynomial(TextEngine &e, const Text &x,
                         const std::vector<double> &coefficients);
Text sum_1d_by_row(TextEngine &e, const Text &x);
fhe_bootstrap_keys rotation_keys_of(const fhe_rotation_key &rk);
void check_rotation_key(const fhe_rotation_key *rk);

// The bootstrap (engine.hip): EvalMod's round, the copy of a result into a
// caller's buffer with its form, the check of an input form, and the
// bootstrap's sequence with the steps that depend on the kind of rotation
// key (a RotationKey or a BootstrapKey).
Text round_text(TextEngine &e, const Text &x);
void copy_out(fhe_context *ctx, TextEngine &e, const Text &t, int64_t *out,
              int *out_level, int *out_flags);
void require_boot_form(const fhe::CkksParams &P, int level, int flags,
                       int min_level, const char *op);
struct BootstrapSteps {
  std::function<void(Text &)> partial_sum; // below full slots, in place
  std::function<std::pair<Text, Text>(const Text &)> coeff_to_slot;
  std::function<Text(const Text &, const Text &)> slot_to_coeff;
};
Text bootstrap_texts(TextEngine &e, const Text &x,
                     const BootstrapSteps &steps);

} // namespace fhe_host
