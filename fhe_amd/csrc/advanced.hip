// The engine's advanced compositions (src/ckks/engine_advanced.cpp) on the
// device: polynomial evaluation through the power basis and the weighted
// sum, and the rotate-and-sum along the slots.  Host C++ over the engine
// operations of text_engine.h, in the reference's operation order, so the
// limbs equal the reference's.
#include "text_engine.h"

namespace fhe_host {

// make_power_basis (src/ckks/engine_advanced.cpp:22-57): x, then the
// squares x^2, x^4, ... by square(ct, rlk), then every other power as the
// product of its highest power-of-two part and the rest; entry p - 1 holds
// x^p.
std::vector<Text> make_power_basis(TextEngine &e, const Text &x,
                                   int max_power) {
  std::vector<Text> basis(size_t(max_power) + 1);
  std::vector<bool> have(size_t(max_power) + 1, false);
  basis[1] = x;
  have[1] = true;
  Text current = x;
  for (int p = 2; p <= max_power; p <<= 1) {
    current = e.multiply(current, current, true);
    basis[p] = current;
    have[p] = true;
  }
  for (int p = 2; p <= max_power; p++) {
    if (have[p])
      continue;
    int high = 1;
    while (high * 2 <= p)
      high *= 2;
    basis[p] = e.multiply(basis[high], basis[p - high], true);
    have[p] = true;
  }
  basis.erase(basis.begin());
  return basis;
}

// evaluate_polynomial (src/ckks/engine_advanced.cpp:365-387) with its
// requirements (:339-362).
Text evaluate_polynomial(TextEngine &e, const Text &x,
                         const std::vector<double> &coefficients) {
  const char *fn = "Evaluate Polynomial With Unit Ciphertext";
  require(x.polys == 2 && !(x.flags & FHE_FLAG_NTT_FORM) &&
              !(x.flags & FHE_FLAG_RESCALED),
          std::string(fn) +
              ": the input should be an INTT-form, not rescaled ciphertext.");
  require(coefficients.size() >= 2,
          std::string(fn) + ": the input coefficient is too short.");
  bool zero_degree = true;
  for (size_t i = 1; i < coefficients.size(); i++)
    zero_degree = zero_degree && coefficients[i] == 0;
  require(!zero_degree, std::string(fn) + ": the input coefficients have "
                                          "only a constant coefficient.");
  const int required =
      int(std::ceil(std::log2(double(coefficients.size() - 1)))) + 1;
  require(x.level >= required,
          std::string(fn) + ": the level of the input is too low.");
  const int max_power = int(coefficients.size()) - 1;
  const std::vector<Text> basis = make_power_basis(e, x, max_power);
  return e.weighted_sum(basis, coefficients);
}

// sum_1d_by_row (src/ckks/engine_advanced.cpp:549-577): the top of the
// one-more-level chain goes one level down, then for delta = 1, 2, 4, ...
// below the slot count the text is rotated by delta and added to itself.
Text sum_1d_by_row(TextEngine &e, const Text &x) {
  require(x.polys == 2 && !(x.flags & FHE_FLAG_NTT_FORM) &&
              !(x.flags & FHE_FLAG_RESCALED),
          "Sum Ciphertext Along Axis 1: the input should be an INTT-form, "
          "not rescaled ciphertext.");
  const int slots = int(e.context()->params.slot_count());
  if (slots == 1)
    return e.clone(x);
  Text current = e.at_top(x) ? e.level_down(x, x.level - 1) : x;
  for (int delta = 1; delta < slots; delta <<= 1) {
    Text rotated = e.rotate(current, delta);
    e.add_inplace(rotated, current);
    current = rotated;
  }
  return current;
}

// A RotationKey as the engine's rotations see it.
fhe_bootstrap_keys rotation_keys_of(const fhe_rotation_key &rk) {
  fhe_bootstrap_keys k{};
  k.rotation_count = rk.count;
  k.rotation = rk.keys;
  k.rotation_moduli_ids = rk.moduli_ids;
  k.rotation_deltas = rk.deltas;
  return k;
}

void check_rotation_key(const fhe_rotation_key *rk) {
  require(rk && rk->count > 0 && rk->keys && rk->moduli_ids && rk->deltas,
          "Rotate: empty rotation key");
  for (int k = 0; k < rk->count; k++)
    require(rk->keys[k] != nullptr, "null argument");
}

void copy_result(fhe_context *ctx, TextEngine &e, const Text &r, int64_t *out,
                 int *out_level, int *out_flags) {
  hip_check(hipMemcpyAsync(out, r.data, size_t(e.words(r)) * 8,
                           hipMemcpyDeviceToDevice, ctx->stream),
            "copy");
  *out_level = r.level;
  *out_flags = r.flags;
}

} // namespace fhe_host

extern "C" {

int fhe_evaluate_polynomial(fhe_context *ctx, int count, const int64_t *in,
                            int level, const double *coefficients,
                            int coefficient_count, const int64_t *rlk,
                            int rlk_moduli_id, int64_t *out, int *out_level,
                            int *out_flags) {
  return guarded([&] {
    require(ctx && in && coefficients && rlk && out && out_level && out_flags,
            "null argument");
    require(count > 0, "Evaluate polynomial: empty batch");
    require(level >= 0 && level <= ctx->params.full_level,
            "Evaluate polynomial: level out of range");
    require(coefficient_count >= 0, "Evaluate polynomial: bad coefficients");
    TextEngine e(ctx, rlk, rlk_moduli_id);
    e.count = count;
    const Text x = e.wrap(in, 2, level, 0);
    const Text r = evaluate_polynomial(
        e, x, std::vector<double>(coefficients,
                                  coefficients + coefficient_count));
    copy_result(ctx, e, r, out, out_level, out_flags);
    ctx->check_launch("Evaluate polynomial");
  });
}

int fhe_sum_1d_by_row(fhe_context *ctx, int count, const int64_t *in,
                      int level, int flags, const fhe_rotation_key *rk,
                      int64_t *out, int *out_level, int *out_flags) {
  return guarded([&] {
    require(ctx && in && out && out_level && out_flags, "null argument");
    require(count > 0, "Sum: empty batch");
    require(level >= 0 && level <= ctx->params.full_level,
            "Sum: level out of range");
    require((flags & ~(FHE_FLAG_NTT_FORM | FHE_FLAG_RESCALED)) == 0,
            "Sum: unknown data flags");
    check_rotation_key(rk);
    const fhe_bootstrap_keys keys = rotation_keys_of(*rk);
    TextEngine e(ctx, nullptr, 0);
    e.count = count;
    e.keys = &keys;
    const Text x = e.wrap(in, 2, level, flags);
    const Text r = sum_1d_by_row(e, x);
    copy_result(ctx, e, r, out, out_level, out_flags);
    ctx->check_launch("Sum");
  });
}

} // extern "C"
