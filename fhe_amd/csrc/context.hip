// C ABI of the context, device memory, transforms and the metric's
// pipeline (fhe_hmult_relin_rescale), declared in include/fhe_amd.h.
#include "context.h"

namespace fhe_host {
thread_local std::string g_last_error;
}

fhe::CkksParams resolve_params(const fhe_params *p, bool twiddles) {
  require(p != nullptr, "null argument");
  require(p->scale_bits == 40 || p->scale_bits == 54,
          "CKKS Parameters: unsupported config.");
  fhe::CkksParams params;
  if (p->prime_count > 0) {
    require(p->primes != nullptr, "primes must be given");
    std::vector<uint64_t> primes(p->primes, p->primes + p->prime_count);
    params = fhe::CkksParams::from_primes(p->log_coeff_count, primes,
                                          p->special_prime_count, p->scale_bits,
                                          p->one_more_level, twiddles);
    if (p->max_level >= 0)
      params.max_level = p->max_level;
  } else {
    require(p->small_prime_bits == 40 || p->small_prime_bits == 54,
            "CKKS Parameters: unsupported config.");
    // The alternating search reproduces the reference's cached chains only
    // up to N = 2^16; the N = 2^17 chain must be passed in explicitly.
    require(p->log_coeff_count <= 16,
            "CKKS Parameters: N = 2^17 needs an explicit prime chain.");
    fhe::CkksConfig c;
    c.log_n = p->log_coeff_count;
    c.special_prime_count = p->special_prime_count;
    c.scale_bits = p->scale_bits;
    c.small_prime_bits = p->small_prime_bits;
    c.one_more_level = p->one_more_level != 0;
    c.max_level = p->max_level;
    params = fhe::CkksParams::from_config(c, twiddles);
  }
  require(params.max_level >= 0 && params.max_level <= params.full_level,
          "The given max level is out of range.");
  return params;
}

// CKKSCore::encode_polynomial (src/ckks/core_encode.cpp:36-150) of `count`
// device messages at `level`, over the level's chain limbs and, with
// include_special, the special primes after them.
void encode_texts(fhe_context *ctx, const double *values, bool complex_input,
                  int count, int level, bool include_special, int64_t *out) {
  require(count > 0, "Encode: empty batch");
  const auto &P = ctx->params;
  require(level >= 0 && level <= P.max_level,
          "Encode: the target level cannot be reached with the current "
          "engine parameters.");
  const auto e = P.encode_level(level);
  const int mid = P.moduli_id_of_level(level);
  const int limbs = P.limbs_at(mid) + (include_special ? P.alpha : 0);
  // the top level of a one-more-level chain leaves limb 0 zero and encodes
  // onto the primes of the level below (src/ckks/core_encode.cpp:64-78)
  const int first = e.top_of_one_more_level ? 1 : 0;
  const int enc_mid =
      e.top_of_one_more_level ? P.moduli_id_of_level(level - 1) : mid;
  require(limbs - first <= fhe::kMaxEwLimbs, "Encode: too many limbs");
  require(int(e.r_mult_scale_after_tile.size()) >= limbs - first,
          "Encode: constants missing");
  require(P.n % fhe::kBlock == 0, "Encode: N below the kernels' block size");
  const auto &tables = ctx->encode_tables();
  void *work = ctx->aux_scratch(fhe::encode_work_bytes(tables, count));
  ctx->stage("encode", [&] {
    fhe::launch_encode(ctx->tables, tables, values, complex_input ? 1 : 0,
                       count, e.scale_before_tile, limbs, first, enc_mid,
                       e.r_mult_scale_after_tile.data(), out, work,
                       ctx->stream);
  });
}

extern "C" {

const char *fhe_last_error(void) { return g_last_error.c_str(); }

int fhe_params_resolve_primes(const fhe_params *p, uint64_t *primes,
                              int capacity, int *prime_count) {
  return guarded([&] {
    require(primes && prime_count, "null argument");
    const auto params = resolve_params(p, false);
    require(capacity >= params.prime_count, "capacity too small");
    std::memcpy(primes, params.q.data(), sizeof(uint64_t) * params.prime_count);
    *prime_count = params.prime_count;
  });
}

int fhe_context_create(const fhe_params *p, fhe_context **out) {
  return guarded([&] {
    require(out != nullptr, "null argument");
    *out = nullptr;
    auto params = resolve_params(p);
    std::unique_ptr<fhe_context> ctx(new fhe_context(std::move(params)));
    // set-up of the allocation-free entry point: this device's address of
    // its scratch array, and the ModUp constants of every ciphertext level
    // with keys at the maximum key-switching level
    noalloc_scratch_setup(ctx->device);
    const auto &P = ctx->params;
    const int key_mid = P.moduli_id_of_level(P.max_key_switching_level());
    for (int m = std::max(key_mid, 0); m < P.chain_count; m++)
      (void)ctx->modup_constants(m, key_mid);
    *out = ctx.release();
  });
}

void fhe_context_destroy(fhe_context *ctx) { delete ctx; }

int fhe_context_info(const fhe_context *ctx, int64_t *info, int count) {
  return guarded([&] {
    require(ctx && info, "null argument");
    const auto &p = ctx->params;
    const int64_t v[FHE_INFO_COUNT] = {p.n,
                                       p.prime_count,
                                       p.chain_count,
                                       p.alpha,
                                       p.full_level,
                                       p.max_level,
                                       p.max_key_switching_level(),
                                       p.n / 2,
                                       p.scale_bits};
    for (int i = 0; i < count && i < FHE_INFO_COUNT; i++)
      info[i] = v[i];
  });
}

int fhe_context_primes(const fhe_context *ctx, uint64_t *out, int capacity) {
  return guarded([&] {
    require(ctx && out, "null argument");
    require(capacity >= ctx->params.prime_count, "capacity too small");
    std::memcpy(out, ctx->params.q.data(),
                sizeof(uint64_t) * ctx->params.prime_count);
  });
}

int fhe_context_prime_constants(const fhe_context *ctx, int i,
                                int64_t *constants3, int64_t *root,
                                int64_t *root_inverse) {
  return guarded([&] {
    require(ctx && constants3, "null argument");
    const auto &p = ctx->params;
    require(i >= 0 && i < p.prime_count, "prime index out of range");
    constants3[0] = int64_t(p.k[i]);
    constants3[1] = p.r2[i];
    constants3[2] = p.ninv[i];
    if (root)
      std::memcpy(root, p.root.data() + size_t(i) * p.n,
                  sizeof(int64_t) * (p.n - 1));
    if (root_inverse)
      std::memcpy(root_inverse, p.root_inverse.data() + size_t(i) * p.n,
                  sizeof(int64_t) * (p.n - 1));
  });
}

void *fhe_context_stream(fhe_context *ctx) {
  return ctx ? static_cast<void *>(ctx->stream) : nullptr;
}

int fhe_context_set_option(fhe_context *ctx, const char *name, int value) {
  return guarded([&] {
    require(ctx && name, "null argument");
    const std::string n(name);
    hip_check(hipStreamSynchronize(ctx->stream), "option sync");
    if (n == "lazy_key_switch")
      ctx->lazy_key_switch = value != 0;
    else if (n == "staged_modup")
      ctx->staged_modup = value != 0;
    else if (n == "fuse_modup")
      ctx->fuse_modup = value != 0;
    else if (n == "fuse_tensor")
      ctx->fuse_tensor = value != 0;
    else if (n == "own_limbs")
      ctx->own_limbs = value != 0;
    else if (n == "overlap_trio")
      ctx->overlap_trio = value < 0 ? 0 : value > 2 ? 2 : value;
    else if (n == "ks_occupancy")
      ctx->ks_occupancy = value;
    else if (n == "ks_prefetch")
      ctx->ks_prefetch = value != 0;
    else if (n == "ks_lazy_accumulate")
      ctx->ks_lazy_accumulate = value != 0;
    else if (n == "ks_wide_accumulate")
      ctx->ks_wide_accumulate = value != 0;
    else if (n == "split_variants")
      ctx->tables.split_variants = value != 0 ? 1 : 0;
    else
      require(false, "unknown option " + n);
  });
}

int fhe_modmul_peak(fhe_context *ctx, int form, double *per_second) {
  return guarded([&] {
    require(ctx && per_second, "null argument");
    require(form == 0 || form == 1, "Modmul peak: unknown form");
    const auto &P = ctx->params;
    // the last small prime of the chain; any twiddle-sized constant
    const int pi = P.chain_count - 1;
    const uint64_t q = P.q[pi], k = P.k[pi];
    const uint64_t w = q / 3;
    const uint64_t wq =
        uint64_t(((unsigned __int128)w << 64) / (unsigned __int128)q);
    int cus = 0;
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                    ctx->device),
              "device attribute");
    const int blocks = std::max(1, cus) * 16;
    void *out = nullptr;
    hip_check(hipMalloc(&out, size_t(blocks) * fhe::kBlock * 8),
              "probe allocation");
    const double rate = fhe::measure_modmul_rate(
        form, blocks, 2048, q, k, w, wq, static_cast<uint64_t *>(out),
        ctx->stream);
    (void)hipFree(out);
    require(rate > 0, "Modmul peak: the probe failed");
    *per_second = rate;
  });
}

int fhe_synchronize(fhe_context *ctx) {
  return guarded([&] {
    require(ctx, "null argument");
    hip_check(hipStreamSynchronize(ctx->stream), "synchronize");
  });
}

int fhe_device_alloc(fhe_context *ctx, size_t bytes, void **out) {
  return guarded([&] {
    require(ctx && out, "null argument");
    *out = nullptr;
    hip_check(hipMalloc(out, std::max<size_t>(bytes, 1)), "device allocation");
  });
}

int fhe_device_free(fhe_context *ctx, void *ptr) {
  return guarded([&] {
    require(ctx, "null argument");
    hip_check(hipStreamSynchronize(ctx->stream), "synchronize");
    hip_check(hipFree(ptr), "device free");
  });
}

int fhe_copy_to_device(fhe_context *ctx, void *dst, const void *src,
                       size_t bytes) {
  return guarded([&] {
    require(ctx && dst && src, "null argument");
    hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice,
                             ctx->stream),
              "copy to device");
    hip_check(hipStreamSynchronize(ctx->stream), "copy to device");
  });
}

int fhe_copy_to_host(fhe_context *ctx, void *dst, const void *src,
                     size_t bytes) {
  return guarded([&] {
    require(ctx && dst && src, "null argument");
    hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost,
                             ctx->stream),
              "copy to host");
    hip_check(hipStreamSynchronize(ctx->stream), "copy to host");
  });
}

int fhe_ntt(fhe_context *ctx, int64_t *data, int polys, int limbs,
            int moduli_id, int enter) {
  return guarded([&] {
    require(ctx && data, "null argument");
    require(polys > 0 && limbs > 0, "NTT: empty shape");
    require(moduli_id >= 0 && moduli_id + limbs <= ctx->params.prime_count,
            "NTT: limbs exceed the prime chain");
    fhe::ntt_forward(ctx->tables, data, data, int64_t(polys) * limbs, limbs,
                     moduli_id, enter ? ctx->tables.r2 : nullptr, ctx->stream);
    ctx->check_launch("NTT");
  });
}

int fhe_intt(fhe_context *ctx, int64_t *data, int polys, int limbs,
             int moduli_id, int post) {
  return guarded([&] {
    require(ctx && data, "null argument");
    require(polys > 0 && limbs > 0, "INTT: empty shape");
    require(post >= 0 && post <= 2, "INTT: unknown epilogue");
    require(moduli_id >= 0 && moduli_id + limbs <= ctx->params.prime_count,
            "INTT: limbs exceed the prime chain");
    fhe::ntt_inverse(ctx->tables, data, data, int64_t(polys) * limbs, limbs,
                     moduli_id, post, ctx->stream);
    ctx->check_launch("INTT");
  });
}

int fhe_mont_mult(fhe_context *ctx, const int64_t *a, const int64_t *b,
                  int64_t *out, int polys, int limbs, int moduli_id) {
  return guarded([&] {
    require(ctx && a && b && out, "null argument");
    require(polys > 0 && limbs > 0, "mont_mult: empty shape");
    require(moduli_id >= 0 && moduli_id + limbs <= ctx->params.prime_count,
            "mont_mult: limbs exceed the prime chain");
    fhe::launch_mont_mult(ctx->tables, a, b, out, int64_t(polys) * limbs,
                          limbs, moduli_id, ctx->stream);
    ctx->check_launch("mont_mult");
  });
}

int fhe_rescale(fhe_context *ctx, const int64_t *in, int64_t *out, int polys,
                int moduli_id) {
  return guarded([&] {
    require(ctx && in && out, "null argument");
    require(polys > 0, "Rescale: empty shape");
    ctx->check_moduli(moduli_id, 2, "Rescale");
    fhe::launch_rescale(ctx->tables, in, out, polys,
                        ctx->params.limbs_at(moduli_id), moduli_id,
                        ctx->stream);
    ctx->check_launch("Rescale");
  });
}

int fhe_relinearize(fhe_context *ctx, int count, int moduli_id,
                    const int64_t *trios, const int64_t *key,
                    int key_moduli_id, int64_t *out) {
  return guarded([&] {
    require(ctx && trios && key && out, "null argument");
    require(count > 0, "Relinearize: empty batch");
    ctx->check_moduli(moduli_id, 1, "Relinearize");
    ctx->check_key(moduli_id, key_moduli_id, "Relinearize");
    ctx->relinearize_into(count, moduli_id, trios, key, key_moduli_id, out);
    ctx->check_launch("Relinearize");
  });
}

} // extern "C"

int64_t hmult_workspace_words(const fhe_context *ctx, int count,
                              int moduli_id) {
  const auto &P = ctx->params;
  const int limbs = P.limbs_at(moduli_id);
  const int ext_limbs = limbs + P.alpha;
  const int64_t duo_l = int64_t(count) * 2 * limbs;
  const int64_t trio_l = int64_t(count) * 3 * limbs;
  const int64_t ext_l = int64_t(count) * ctx->ext_slots(moduli_id) * ext_limbs;
  const int64_t acc_l = int64_t(count) * 2 * ext_limbs;
  return (2 * duo_l + 2 * trio_l + ext_l + acc_l) * int64_t(P.n);
}

// The batched HMult + relinearize + rescale pipeline on the workspace ws
// (hmult_workspace_words) and, when given, the digits buffer (the staged
// ModUp's Garner digits; else the context's second buffer).  Arguments are
// checked by the caller.
void hmult_relin_rescale_into(fhe_context *ctx, int count, int moduli_id,
                              const int64_t *a, const int64_t *b,
                              const int64_t *key, int key_moduli_id,
                              int64_t *out, int64_t *ws, int64_t *digits) {
    const auto &P = ctx->params;
    const int limbs = P.limbs_at(moduli_id);
    const int ext_limbs = limbs + P.alpha;
    const int64_t N = ctx->n();
    const int64_t duo_l = int64_t(count) * 2 * limbs;
    const int64_t trio_l = int64_t(count) * 3 * limbs;
    const int64_t ext_l =
        int64_t(count) * ctx->ext_slots(moduli_id) * ext_limbs;
    // trio_c: the coefficient form of the trio; trio keeps the NTT form,
    // whose d2 limbs serve the partitions' own extended limbs
    int64_t *an = ws, *bn = an + duo_l * N, *trio = bn + duo_l * N;
    int64_t *trio_c = trio + trio_l * N;
    int64_t *ext = trio_c + trio_l * N, *acc = ext + ext_l * N;
    // prepare_for_multiply: to NTT form (src/ckks/engine_multiply.cpp:13-27)
    // Only a1 * b1 reaches the representative-sensitive ModUp (through the
    // exact INTT of d2), so the transforms of a0 and b0 may be lazy.
    const fhe::LazySpec first_poly =
        ctx->lazy_key_switch ? fhe::LazySpec{2, 1} : fhe::LazySpec();
    // relinearize: to_intt of all three, key switch of the third; d0 and d1
    // are canonicalized by the relinearization add, d2 feeds ModUp exactly
    const fhe::LazySpec d01 =
        ctx->lazy_key_switch ? fhe::LazySpec{3, 2} : fhe::LazySpec();
    const bool overlap =
        ctx->overlap_trio && ctx->lazy_key_switch &&
        fhe::ntt_parts_supported(ctx->tables) &&
        !(ctx->fuse_tensor && fhe::tensor_fused_supported(ctx->tables));
    // once the side stream has forked, every exit (an exception included)
    // makes the main stream wait for it, so later work on the workspace is
    // ordered after the side-stream INTT
    bool joined = true;
    struct Join {
      fhe_context *ctx;
      bool &joined;
      ~Join() {
        if (!joined)
          (void)hipStreamWaitEvent(ctx->stream, ctx->join_event, 0);
      }
    } join_guard{ctx, joined};
    if (ctx->fuse_tensor && fhe::tensor_fused_supported(ctx->tables)) {
      ctx->stage("ntt_pass_high(inputs)", [&] {
        fhe::ntt_forward_first(ctx->tables, a, an, duo_l, limbs, moduli_id,
                               ctx->tables.r2, ctx->stream, first_poly);
        fhe::ntt_forward_first(ctx->tables, b, bn, duo_l, limbs, moduli_id,
                               ctx->tables.r2, ctx->stream, first_poly);
      });
      ctx->stage("tensor_r16", [&] {
        fhe::launch_tensor_fused(ctx->tables, an, bn, trio, count, limbs,
                                 moduli_id, ctx->lazy_key_switch,
                                 ctx->stream);
      });
      ctx->stage("intt_pass_high(trio)", [&] {
        fhe::ntt_inverse_second(ctx->tables, trio, trio, trio_l, limbs,
                                moduli_id, fhe::kInttExit, ctx->stream, d01);
      });
    } else {
      ctx->stage("ntt(inputs)", [&] {
        fhe::ntt_forward(ctx->tables, a, an, duo_l, limbs, moduli_id,
                         ctx->tables.r2, ctx->stream, first_poly);
        fhe::ntt_forward(ctx->tables, b, bn, duo_l, limbs, moduli_id,
                         ctx->tables.r2, ctx->stream, first_poly);
      });
      ctx->stage("tensor", [&] {
        fhe::launch_tensor(ctx->tables, an, bn, trio, count, limbs, moduli_id,
                           ctx->stream);
      });
      if (overlap) {
        // d0, d1 (lazy) on the side stream, forked once the trio is written
        // (overlap_trio 1) or once d2's INTT is done (2, alongside the key
        // switching of d2); d2 (exact, feeds ModUp) here
        fhe::LazySpec lazy_part = d01, exact_part = d01;
        lazy_part.part = 2;
        exact_part.part = 1;
        auto fork = [&] {
          hip_check(hipEventRecord(ctx->fork_event, ctx->stream), "event");
          hip_check(hipStreamWaitEvent(ctx->side, ctx->fork_event, 0),
                    "event");
          fhe::ntt_inverse(ctx->tables, trio, trio_c, trio_l, limbs,
                           moduli_id, fhe::kInttExit, ctx->side, lazy_part);
          hip_check(hipEventRecord(ctx->join_event, ctx->side), "event");
          joined = false;
        };
        if (ctx->overlap_trio == 1)
          fork();
        ctx->stage("intt(trio)", [&] {
          fhe::ntt_inverse(ctx->tables, trio, trio_c, trio_l, limbs,
                           moduli_id, fhe::kInttExit, ctx->stream,
                           exact_part);
        });
        if (ctx->overlap_trio != 1)
          fork();
      } else {
        ctx->stage("intt(trio)", [&] {
          fhe::ntt_inverse(ctx->tables, trio, trio_c, trio_l, limbs,
                           moduli_id, fhe::kInttExit, ctx->stream, d01);
        });
      }
    }
    const bool fused_trio = ctx->fuse_tensor &&
                            fhe::tensor_fused_supported(ctx->tables);
    const int64_t *coef = fused_trio ? trio : trio_c;
    ctx->key_switch(count, moduli_id, key, key_moduli_id, coef + 2 * limbs * N,
                    3 * limbs * N, ext, acc,
                    fused_trio ? nullptr : trio + 2 * limbs * N,
                    3 * limbs * N, false, digits);
    if (overlap) {
      joined = true;
      hip_check(hipStreamWaitEvent(ctx->stream, ctx->join_event, 0), "event");
    }
    ctx->stage("moddown_relin_rescale", [&] {
      fhe::launch_moddown_relin_rescale(ctx->tables, acc, coef, out, count,
                                        limbs, moduli_id, ctx->stream);
    });
}

extern "C" {

int fhe_hmult_relin_rescale(fhe_context *ctx, int count, int moduli_id,
                            const int64_t *a, const int64_t *b,
                            const int64_t *key, int key_moduli_id,
                            int64_t *out) {
  return guarded([&] {
    require(ctx && a && b && key && out, "null argument");
    require(count > 0, "Multiply: empty batch");
    ctx->check_moduli(moduli_id, 3, "Multiply");
    ctx->check_key(moduli_id, key_moduli_id, "Multiply");
    int64_t *ws = ctx->scratch(
        size_t(hmult_workspace_words(ctx, count, moduli_id)) * 8);
    hmult_relin_rescale_into(ctx, count, moduli_id, a, b, key, key_moduli_id,
                             out, ws, nullptr);
    ctx->check_launch("Multiply");
  });
}

static void require_text_form(const fhe::CkksParams &P, int level, int flags,
                              const char *op) {
  require((flags & ~(FHE_FLAG_NTT_FORM | FHE_FLAG_RESCALED)) == 0,
          std::string(op) + ": unknown data flags");
  // Requirement::POSITIVE_LEVEL (src/ckks/engine_require.cpp:495-500)
  require(level >= 1 && level <= P.full_level,
          std::string(op) + ": the text should have a positive level.");
}

// Shared body of fhe_multiply and fhe_multiply_plain: both inputs prepared
// at the lower level, then the tensor (b with two polynomials) or the
// plaintext product (b with one), then the optional relinearization.
static void multiply_texts(fhe_context *ctx, int count, const int64_t *a,
                           int a_polys, int a_level, int a_flags,
                           const int64_t *b, int b_polys, int b_level,
                           int b_flags, const int64_t *key, int key_moduli_id,
                           int64_t *out, const char *op) {
  const auto &P = ctx->params;
  require_text_form(P, a_level, a_flags, op);
  require_text_form(P, b_level, b_flags, op);
  const int level = std::min(a_level, b_level);
  const int mid = ctx->moduli_id_of(level, true);
  const int limbs = P.limbs_at(mid);
  const int64_t N = ctx->n();
  const int64_t poly = int64_t(limbs) * N;
  const int prod_polys = b_polys == 2 ? 3 : a_polys;
  if (key)
    ctx->check_key(mid, key_moduli_id, op);
  // prepared inputs and the product in the stash; the preparation's
  // temporaries in the scratch workspace, which relinearization reuses
  const int64_t an_l = int64_t(count) * a_polys * poly;
  const int64_t bn_l = int64_t(count) * b_polys * poly;
  const int64_t prod_l = int64_t(count) * prod_polys * poly;
  int64_t *an = ctx->stash(size_t(an_l + bn_l + (key ? prod_l : 0)) * 8);
  int64_t *bn = an + an_l;
  int64_t *prod = key ? bn + bn_l : out;
  auto limbs_of = [&](int l, int f) {
    return int64_t(P.limbs_at(ctx->moduli_id_of(l, f & FHE_FLAG_RESCALED)));
  };
  const int64_t t_l =
      std::max(int64_t(count) * a_polys * limbs_of(a_level, a_flags),
               int64_t(count) * b_polys * limbs_of(b_level, b_flags)) *
      N;
  int64_t *t0 = ctx->scratch(size_t(2 * t_l) * 8), *t1 = t0 + t_l;
  const int64_t *ap = ctx->prepare_for_multiply_into(
      a, int64_t(count) * a_polys, a_level, a_flags, level, an, t0, t1);
  const int64_t *bp = ctx->prepare_for_multiply_into(
      b, int64_t(count) * b_polys, b_level, b_flags, level, bn, t0, t1);
  if (b_polys == 2) {
    // CKKSCore::multiply / square (src/ckks/core.cpp:284-343)
    ctx->stage("tensor", [&] {
      fhe::launch_tensor(ctx->tables, ap, bp, prod, count, limbs, mid,
                         ctx->stream);
    });
  } else {
    // CKKSCore::multiply(ct, pt) (src/ckks/core.cpp:257-282)
    ctx->stage("mont_mult", [&] {
      for (int c = 0; c < count; c++)
        for (int p = 0; p < a_polys; p++)
          fhe::launch_mont_mult(ctx->tables, ap + (c * a_polys + p) * poly,
                                bp + c * poly,
                                prod + (c * prod_polys + p) * poly, limbs,
                                limbs, mid, ctx->stream);
    });
  }
  if (key)
    ctx->relinearize_into(count, mid, prod, key, key_moduli_id, out);
}

int fhe_multiply(fhe_context *ctx, int count, const int64_t *a, int a_level,
                 int a_flags, const int64_t *b, int b_level, int b_flags,
                 const int64_t *key, int key_moduli_id, int64_t *out) {
  return guarded([&] {
    require(ctx && a && b && out, "null argument");
    require(count > 0, "Multiply: empty batch");
    multiply_texts(ctx, count, a, 2, a_level, a_flags, b, 2, b_level, b_flags,
                   key, key_moduli_id, out, "Multiply");
    ctx->check_launch("Multiply");
  });
}

int fhe_multiply_plain(fhe_context *ctx, int count, const int64_t *ct,
                       int ct_polys, int ct_level, int ct_flags,
                       const int64_t *pt, int pt_level, int pt_flags,
                       int64_t *out) {
  return guarded([&] {
    require(ctx && ct && pt && out, "null argument");
    require(count > 0, "Multiply: empty batch");
    require(ct_polys == 2 || ct_polys == 3,
            "Multiply: the ciphertext should be a duo or a trio.");
    multiply_texts(ctx, count, ct, ct_polys, ct_level, ct_flags, pt, 1,
                   pt_level, pt_flags, nullptr, 0, out, "Multiply");
    ctx->check_launch("Multiply");
  });
}

int fhe_add_texts(fhe_context *ctx, int count, const int64_t *a,
                  int a_polys, int a_level, int a_flags, const int64_t *b,
                  int b_polys, int b_level, int b_flags, int subtract,
                  int64_t *out) {
  return guarded([&] {
    require(ctx && a && b && out, "null argument");
    require(count > 0, "Add: empty batch");
    require((a_polys == 2 || a_polys == 3) && (b_polys == 2 || b_polys == 3),
            "Add: the ciphertexts should be duos or trios.");
    const auto &P = ctx->params;
    for (int f : {a_flags, b_flags})
      require((f & ~(FHE_FLAG_NTT_FORM | FHE_FLAG_RESCALED)) == 0,
              "Add: unknown data flags");
    require(a_level >= 0 && a_level <= P.full_level && b_level >= 0 &&
                b_level <= P.full_level,
            "Add: level out of range");
    // CKKSEngine::match_form (src/ckks/engine_add.cpp:37-74): the first
    // text's form decides the domain; the higher text leaves the NTT domain
    // before it goes down; a rescale where only one of them is rescaled
    const bool ntt_form = (a_flags & FHE_FLAG_NTT_FORM) != 0;
    const int level = std::min(a_level, b_level);
    const bool ra = (a_flags & FHE_FLAG_RESCALED) && a_level == level;
    const bool rb = (b_flags & FHE_FLAG_RESCALED) && b_level == level;
    const bool rescaled = ra || rb;
    const int mid = ctx->moduli_id_of(level, rescaled);
    const int limbs = P.limbs_at(mid);
    const int64_t N = ctx->n();
    const int64_t poly = int64_t(limbs) * N;
    auto limbs_of = [&](int l, int f) {
      return int64_t(P.limbs_at(ctx->moduli_id_of(l, f & FHE_FLAG_RESCALED)));
    };
    const int64_t an_l = int64_t(count) * a_polys * poly;
    const int64_t bn_l = int64_t(count) * b_polys * poly;
    int64_t *an = ctx->stash(size_t(an_l + bn_l) * 8), *bn = an + an_l;
    const int64_t t_l =
        std::max(int64_t(count) * a_polys * limbs_of(a_level, a_flags),
                 int64_t(count) * b_polys * limbs_of(b_level, b_flags)) *
        N;
    int64_t *t0 = ctx->scratch(size_t(2 * t_l) * 8), *t1 = t0 + t_l;
    const int64_t *ap = ctx->transform_into(
        a, int64_t(count) * a_polys, a_level, a_flags,
        !ntt_form || a_level > b_level, level, rescaled, ntt_form, an, t0, t1);
    const int64_t *bp = ctx->transform_into(
        b, int64_t(count) * b_polys, b_level, b_flags,
        !ntt_form || b_level > a_level, level, rescaled, ntt_form, bn, t0, t1);
    const int op = subtract ? fhe::kEwSubReduce : fhe::kEwAddReduce;
    ctx->stage("add", [&] {
      if (a_polys == b_polys) {
        // CKKSCore::add / subtract (src/ckks/core.cpp:415-435, 598-614)
        fhe::launch_elementwise(ctx->tables, op, ap, bp, out,
                                int64_t(count) * a_polys * limbs, limbs, mid,
                                nullptr, ctx->stream);
        return;
      }
      // add_trio_and_duo, subtract_trio_by_duo, subtract_duo_by_trio
      // (src/ckks/core.cpp:437-460, 615-660): the first two polynomials
      // combined, the trio's third copied, or reduced and negated when the
      // trio is subtracted
      const bool a_trio = a_polys == 3;
      for (int c = 0; c < count; c++) {
        const int64_t *x1 = ap + int64_t(c) * a_polys * poly;
        const int64_t *x2 = bp + int64_t(c) * b_polys * poly;
        int64_t *y = out + int64_t(c) * 3 * poly;
        if (!subtract && !a_trio)
          std::swap(x1, x2);
        fhe::launch_elementwise(ctx->tables, op, x1, x2, y, 2 * limbs, limbs,
                                mid, nullptr, ctx->stream);
        const int64_t *third = (a_trio ? ap + int64_t(c) * 3 * poly
                                       : bp + int64_t(c) * 3 * poly) +
                               2 * poly;
        if (subtract && !a_trio) {
          fhe::launch_elementwise(ctx->tables, fhe::kEwReduce, third, nullptr,
                                  y + 2 * poly, limbs, limbs, mid, nullptr,
                                  ctx->stream);
          fhe::launch_elementwise(ctx->tables, fhe::kEwNegateMakeUnsigned,
                                  y + 2 * poly, nullptr, y + 2 * poly, limbs,
                                  limbs, mid, nullptr, ctx->stream);
        } else {
          hip_check(hipMemcpyAsync(y + 2 * poly, third, size_t(poly) * 8,
                                   hipMemcpyDeviceToDevice, ctx->stream),
                    "copy");
        }
      }
    });
    ctx->check_launch("Add");
  });
}

static void multiplier_of(const fhe::CkksParams &P, int kind, double value,
                          int64_t int_value, int level, int level2,
                          int64_t *out, int capacity, int *count) {
  {
    auto level_ok = [&](int l) { return l >= 0 && l <= P.full_level; };
    std::vector<int64_t> v;
    switch (kind) {
    case FHE_MULTIPLIER_ADD:
      require(level_ok(level), "multiplier: level out of range");
      v = P.add_multiplier(value, level);
      break;
    case FHE_MULTIPLIER_MULT:
      require(level_ok(level) && level >= 1, "multiplier: level out of range");
      v = P.mult_multiplier(value, level);
      break;
    case FHE_MULTIPLIER_MULT_INT:
      require(level >= 0 && level < P.chain_count,
              "multiplier: moduli_id out of range");
      v = P.mult_multiplier_int(int_value, level);
      break;
    case FHE_MULTIPLIER_LEVEL_DOWN:
      require(level_ok(level) && level_ok(level2) && level2 < level,
              "multiplier: levels out of range");
      v = P.level_down_multiplier(level, level2);
      break;
    case FHE_MULTIPLIER_LEVEL_DOWN_AFTER_RESCALE:
      require(level_ok(level) && level >= 1 && level_ok(level2) &&
                  level2 < level,
              "multiplier: levels out of range");
      v = P.level_down_multiplier_after_rescale(level, level2);
      break;
    default:
      require(false, "multiplier: unknown kind");
    }
    require(capacity >= int(v.size()), "multiplier: capacity too small");
    std::memcpy(out, v.data(), v.size() * sizeof(int64_t));
    *count = int(v.size());
  }
}

int fhe_multiplier(const fhe_context *ctx, int kind, double value,
                   int64_t int_value, int level, int level2, int64_t *out,
                   int capacity, int *count) {
  return guarded([&] {
    require(ctx && out && count, "null argument");
    multiplier_of(ctx->params, kind, value, int_value, level, level2, out,
                  capacity, count);
  });
}

int fhe_params_multiplier(const fhe_params *p, int kind, double value,
                          int64_t int_value, int level, int level2,
                          int64_t *out, int capacity, int *count) {
  return guarded([&] {
    require(out && count, "null argument");
    multiplier_of(resolve_params(p, false), kind, value, int_value, level,
                  level2, out, capacity, count);
  });
}

int fhe_params_encode_constants(const fhe_params *p, int level,
                                long double *ckks_scale,
                                double *scale_before_tile,
                                int64_t *log_scale_after_tile,
                                int *top_of_one_more_level, int64_t *r_mult,
                                int capacity, int *count) {
  return guarded([&] {
    require(ckks_scale && scale_before_tile && log_scale_after_tile &&
                top_of_one_more_level && r_mult && count,
            "null argument");
    const auto P = resolve_params(p, false);
    require(level >= 0 && level <= P.full_level,
            "encode constants: level out of range");
    const auto e = P.encode_level(level);
    const auto &v = e.r_mult_scale_after_tile;
    require(capacity >= int(v.size()), "encode constants: capacity too small");
    *ckks_scale = e.ckks_scale;
    *scale_before_tile = e.scale_before_tile;
    *log_scale_after_tile = e.log_scale_after_tile;
    *top_of_one_more_level = e.top_of_one_more_level ? 1 : 0;
    std::memcpy(r_mult, v.data(), v.size() * sizeof(int64_t));
    *count = int(v.size());
  });
}

int fhe_encode(fhe_context *ctx, const double *values, int complex_input,
               int count, int level, int64_t *out) {
  return guarded([&] {
    require(ctx && values && out, "null argument");
    encode_texts(ctx, values, complex_input != 0, count, level, false, out);
    ctx->check_launch("Encode");
  });
}

int fhe_decode(fhe_context *ctx, const int64_t *pt, int count, int moduli_id,
               long double ckks_scale, int complex_output, double *out) {
  return guarded([&] {
    require(ctx && pt && out, "null argument");
    require(count > 0, "Decode: empty batch");
    const auto &P = ctx->params;
    require(moduli_id >= 0 && moduli_id < P.chain_count,
            "Decode: moduli_id out of range.");
    // CKKSCore::convert_plaintext_to_double (src/ckks/core_decode.cpp:10-107)
    require(P.limbs_at(moduli_id) > 1,
            "A plaintext needs at least two chains to be decoded.");
    long double modulus = 1;
    for (int i = 0; i < P.limbs_at(moduli_id); i++)
      modulus *= static_cast<long double>(P.q[moduli_id + i]);
    require(!(ckks_scale > modulus),
            "Overflow due to large ckks_scale and small unit plaintext "
            "modulus.");
    // rescale while the rescaled scale stays above the base scale
    int mid = moduli_id;
    long double scale = ckks_scale;
    const int64_t N = ctx->n();
    const int64_t *src = pt;
    int64_t *bufs = nullptr;
    int which = 0;
    while (mid <= P.full_level) {
      const long double rescaled = scale / P.q[mid];
      if (rescaled <= P.scale)
        break;
      if (!bufs)
        bufs = ctx->scratch(size_t(2) * count * P.limbs_at(moduli_id) * N * 8);
      int64_t *dst = bufs + int64_t(which) * count * P.limbs_at(moduli_id) * N;
      ctx->stage("decode(rescale)", [&] {
        fhe::launch_rescale(ctx->tables, src, dst, count, P.limbs_at(mid), mid,
                            ctx->stream);
      });
      src = dst;
      which ^= 1;
      scale = rescaled;
      mid++;
    }
    const double scale_inverse = static_cast<double>(1.0 / scale);
    const auto &d = ctx->decode_tables();
    const auto &crt = ctx->decode_crt(mid);
    void *work = ctx->aux_scratch(fhe::decode_work_bytes(d, count));
    ctx->stage("decode", [&] {
      fhe::launch_decode(ctx->tables, d, crt, src, count, P.limbs_at(mid), mid,
                         scale_inverse, complex_output ? 1 : 0, out, work,
                         ctx->stream);
    });
    ctx->check_launch("Decode");
  });
}

int fhe_decoded_scale(const fhe_context *ctx, int moduli_id,
                      long double ckks_scale, int *decoded_moduli_id,
                      long double *decoded_scale) {
  return guarded([&] {
    require(ctx && decoded_moduli_id && decoded_scale, "null argument");
    const auto &P = ctx->params;
    require(moduli_id >= 0 && moduli_id < P.chain_count,
            "Decode: moduli_id out of range.");
    int mid = moduli_id;
    long double scale = ckks_scale;
    while (mid <= P.full_level) {
      const long double rescaled = scale / P.q[mid];
      if (rescaled <= P.scale)
        break;
      scale = rescaled;
      mid++;
    }
    *decoded_moduli_id = mid;
    *decoded_scale = scale;
  });
}

int fhe_elementwise(fhe_context *ctx, int op, const int64_t *x1,
                    const int64_t *x2, int64_t *y, int polys, int limbs,
                    int moduli_id, const int64_t *constants) {
  return guarded([&] {
    require(ctx && x1 && y, "null argument");
    require(op >= 0 && op < fhe::kEwOpCount, "elementwise: unknown operation");
    require(polys > 0 && limbs > 0, "elementwise: empty shape");
    require(moduli_id >= 0 && moduli_id + limbs <= ctx->params.prime_count,
            "elementwise: limbs exceed the prime chain");
    require(limbs <= fhe::kMaxEwLimbs, "elementwise: too many limbs");
    require(!fhe::ew_needs_x2(op) || x2 != nullptr,
            "elementwise: the operation needs a second operand");
    require(!fhe::ew_needs_constants(op) || constants != nullptr,
            "elementwise: the operation needs per-limb constants");
    ctx->stage("elementwise", [&] {
      fhe::launch_elementwise(ctx->tables, op, x1, x2, y,
                              int64_t(polys) * limbs, limbs, moduli_id,
                              constants, ctx->stream);
    });
    ctx->check_launch("elementwise");
  });
}

// ---- scalar operations on INTT-form, not-rescaled data at a level

int fhe_multiply_double(fhe_context *ctx, const int64_t *in, int64_t *out,
                        int polys, int level, double value) {
  return guarded([&] {
    require(ctx && in && out, "null argument");
    require(polys > 0, "Multiply: empty shape");
    const auto &P = ctx->params;
    require(level >= 1 && level <= P.full_level,
            "Multiply: the level must be positive.");
    const int mid = P.moduli_id_of_level(level);
    const auto c = P.mult_multiplier(value, level);
    require(int(c.size()) <= fhe::kMaxEwLimbs, "Multiply: too many limbs");
    ctx->stage("scalar_rescale", [&] {
      fhe::launch_scalar_rescale(ctx->tables, in, out, polys, int(c.size()),
                                 mid, c.data(), 0, ctx->stream);
    });
    ctx->check_launch("Multiply");
  });
}

int fhe_multiply_int(fhe_context *ctx, const int64_t *in, int64_t *out,
                     int polys, int moduli_id, int64_t value) {
  return guarded([&] {
    require(ctx && in && out, "null argument");
    require(polys > 0, "Multiply: empty shape");
    ctx->check_moduli(moduli_id, 1, "Multiply");
    const auto c = ctx->params.mult_multiplier_int(value, moduli_id);
    require(int(c.size()) <= fhe::kMaxEwLimbs, "Multiply: too many limbs");
    ctx->stage("elementwise", [&] {
      fhe::launch_elementwise(ctx->tables, fhe::kEwEnterReduce, in, nullptr,
                              out, int64_t(polys) * int64_t(c.size()),
                              int(c.size()), moduli_id, c.data(),
                              ctx->stream);
    });
    ctx->check_launch("Multiply");
  });
}

int fhe_add_double(fhe_context *ctx, const int64_t *in, int64_t *out,
                   int count, int polys, int level, double value) {
  return guarded([&] {
    require(ctx && in && out, "null argument");
    require(count > 0 && polys > 0, "Add: empty shape");
    const auto &P = ctx->params;
    require(level >= 0 && level <= P.full_level, "Add: level out of range");
    const int mid = P.moduli_id_of_level(level);
    const auto c = P.add_multiplier(value, level);
    const int limbs = int(c.size());
    require(limbs <= fhe::kMaxEwLimbs, "Add: too many limbs");
    ctx->stage("elementwise", [&] {
      // polynomial 0 of every ciphertext takes the scalar, the others are
      // copied: one launch for the batch
      fhe::launch_elementwise(ctx->tables, fhe::kEwAddScalarReduce, in,
                              nullptr, out, int64_t(count) * polys * limbs,
                              limbs, mid, c.data(), ctx->stream, polys);
    });
    ctx->check_launch("Add");
  });
}

int fhe_level_down(fhe_context *ctx, const int64_t *in, int64_t *out,
                   int polys, int level, int rescaled, int dst_level) {
  return guarded([&] {
    require(ctx && in && out, "null argument");
    require(polys > 0, "Level down: empty shape");
    const auto &P = ctx->params;
    require(level >= 0 && level <= P.full_level && dst_level >= 0 &&
                dst_level < level,
            "Level down: the destination level must be below the level.");
    ctx->level_down_into(in, out, polys, level, rescaled != 0, dst_level);
    ctx->check_launch("Level down");
  });
}

int fhe_encrypt_pk(fhe_context *ctx, int count, int moduli_id,
                   const int64_t *pt, const int64_t *pk, int pk_moduli_id,
                   const int64_t *u, const int64_t *e0, const int64_t *e1,
                   int64_t *out) {
  return guarded([&] {
    require(ctx && pt && pk && u && e0 && e1 && out, "null argument");
    require(count > 0, "Encrypt: empty batch");
    ctx->check_moduli(moduli_id, 1, "Encrypt");
    require(pk_moduli_id >= 0 && pk_moduli_id <= moduli_id,
            "Encrypt: the key level is lower than the plaintext level.");
    const auto &P = ctx->params;
    const int limbs = P.limbs_at(moduli_id);
    const int pk_limbs = P.limbs_at(pk_moduli_id);
    const int64_t N = ctx->n();
    const int64_t poly = int64_t(limbs) * N;
    int64_t *un = ctx->scratch(size_t(count) * poly * 8);
    ctx->stage("encrypt", [&] {
      fhe::launch_tile_unsigned(ctx->tables, u, un, count, limbs, moduli_id,
                                ctx->stream);
      fhe::ntt_forward(ctx->tables, un, un, int64_t(count) * limbs, limbs,
                       moduli_id, ctx->tables.r2, ctx->stream);
      for (int c = 0; c < count; c++)
        for (int h = 0; h < 2; h++)
          fhe::launch_mont_mult(
              ctx->tables, un + c * poly,
              pk + (int64_t(h) * pk_limbs + (moduli_id - pk_moduli_id)) * N,
              out + (2 * int64_t(c) + h) * poly, limbs, limbs, moduli_id,
              ctx->stream);
      fhe::ntt_inverse(ctx->tables, out, out, int64_t(count) * 2 * limbs,
                       limbs, moduli_id, fhe::kInttExit, ctx->stream);
      fhe::launch_encrypt_finish(ctx->tables, pt, e0, e1, out, count, limbs,
                                 moduli_id, ctx->stream);
    });
    ctx->check_launch("Encrypt");
  });
}

int fhe_decrypt(fhe_context *ctx, int count, int moduli_id, const int64_t *ct,
                const int64_t *sk, int sk_moduli_id, int64_t *out) {
  return guarded([&] {
    require(ctx && ct && sk && out, "null argument");
    require(count > 0, "Decrypt: empty batch");
    ctx->check_moduli(moduli_id, 1, "Decrypt");
    require(sk_moduli_id >= 0 && sk_moduli_id <= moduli_id,
            "Decrypt: the key level is lower than the ciphertext level.");
    require(static_cast<const void *>(out) != static_cast<const void *>(ct),
            "Decrypt: the output must not alias the ciphertext");
    const auto &P = ctx->params;
    const int limbs = P.limbs_at(moduli_id);
    const int64_t N = ctx->n();
    const int64_t poly = int64_t(limbs) * N;
    const int64_t *s = sk + int64_t(moduli_id - sk_moduli_id) * N;
    ctx->stage("decrypt", [&] {
      // out_c = to_intt(to_ntt(c1) * s) (src/ckks/core.cpp:171-194)
      for (int c = 0; c < count; c++) {
        fhe::ntt_forward(ctx->tables, ct + (2 * int64_t(c) + 1) * poly,
                         out + c * poly, limbs, limbs, moduli_id,
                         ctx->tables.r2, ctx->stream);
        fhe::launch_mont_mult(ctx->tables, out + c * poly, s, out + c * poly,
                              limbs, limbs, moduli_id, ctx->stream);
      }
      fhe::ntt_inverse(ctx->tables, out, out, int64_t(count) * limbs, limbs,
                       moduli_id, fhe::kInttExit, ctx->stream);
      for (int c = 0; c < count; c++)
        fhe::launch_elementwise(ctx->tables, fhe::kEwAddReduce,
                                ct + 2 * int64_t(c) * poly, out + c * poly,
                                out + c * poly, limbs, limbs, moduli_id,
                                nullptr, ctx->stream);
    });
    ctx->check_launch("Decrypt");
  });
}

int fhe_profile_enable(fhe_context *ctx, int on) {
  return guarded([&] {
    require(ctx != nullptr, "null argument");
    ctx->require_not_capturing("Stage timing");
    ctx->fold_timings();
    if (on) {
      ctx->stage_names.clear();
      ctx->stage_ms.clear();
      ctx->stage_launches.clear();
    }
    ctx->timing = on != 0;
  });
}

int fhe_profile_entry(fhe_context *ctx, int index, char *name,
                      int name_capacity, double *total_ms,
                      int64_t *launches) {
  return guarded([&] {
    require(ctx && name && total_ms && launches && name_capacity > 0,
            "null argument");
    ctx->fold_timings();
    require(index >= 0 && index < int(ctx->stage_names.size()),
            "profile: index out of range");
    std::snprintf(name, size_t(name_capacity), "%s",
                  ctx->stage_names[index].c_str());
    *total_ms = ctx->stage_ms[index];
    *launches = ctx->stage_launches[index];
  });
}

int fhe_profile_count(fhe_context *ctx, int *count) {
  return guarded([&] {
    require(ctx && count, "null argument");
    ctx->fold_timings();
    *count = int(ctx->stage_names.size());
  });
}

int fhe_automorphism(fhe_context *ctx, const int64_t *in, int64_t *out,
                     int polys, int limbs, int moduli_id, int delta,
                     int ntt_form) {
  return guarded([&] {
    require(ctx && in && out, "null argument");
    require(polys > 0 && limbs > 0, "Rotate: empty shape");
    require(moduli_id >= 0 && moduli_id + limbs <= ctx->params.prime_count,
            "Rotate: limbs exceed the prime chain");
    require(in != out, "Rotate: the output must not alias the input");
    const uint32_t g = automorphism_exponent(ctx->params.log_n, delta);
    if (ntt_form)
      fhe::launch_ntt_automorphism(ctx->tables, in, out,
                                   int64_t(polys) * limbs, g, ctx->stream);
    else
      fhe::launch_automorphism(ctx->tables, in, out, int64_t(polys) * limbs,
                               limbs, moduli_id, g, ctx->stream);
    ctx->check_launch("Rotate");
  });
}

int fhe_rotate(fhe_context *ctx, int count, int moduli_id, const int64_t *in,
               const int64_t *key, int key_moduli_id, int delta,
               int64_t *out) {
  return guarded([&] {
    require(ctx && in && key && out, "null argument");
    require(count > 0, "Rotate: empty batch");
    ctx->check_moduli(moduli_id, 1, "Rotate");
    ctx->check_key(moduli_id, key_moduli_id, "Rotate");
    const auto &P = ctx->params;
    const int limbs = P.limbs_at(moduli_id);
    const int ext_limbs = limbs + P.alpha;
    const int64_t N = ctx->n();
    const int64_t duo_l = int64_t(count) * 2 * limbs;
    const int64_t ext_l =
        int64_t(count) * ctx->ext_slots(moduli_id) * ext_limbs;
    const int64_t acc_l = int64_t(count) * 2 * ext_limbs;
    int64_t *ws = ctx->scratch(size_t(duo_l + ext_l + acc_l) * N * 8);
    int64_t *sw = ws, *ext = sw + duo_l * N, *acc = ext + ext_l * N;
    // key switch of c1; c0 joins after ModDown, where the reference adds
    // c0 * P before it (src/ckks/core_key_switch.cpp:509-548): both give
    // floor((X + c0 P) / P) = floor(X / P) + c0
    ctx->key_switch(count, moduli_id, key, key_moduli_id, in + limbs * N,
                    2 * limbs * N, ext, acc);
    ctx->stage("moddown_add", [&] {
      fhe::launch_moddown_add(ctx->tables, acc, in, 2 * limbs * N, 1, sw,
                              count, limbs, moduli_id, ctx->stream);
    });
    ctx->stage("automorphism", [&] {
      fhe::launch_automorphism(ctx->tables, sw, out, duo_l, limbs, moduli_id,
                               automorphism_exponent(P.log_n, delta),
                               ctx->stream);
    });
    ctx->check_launch("Rotate");
  });
}

int fhe_rotate_batch(fhe_context *ctx, int count, int moduli_id,
                     const int64_t *in, int key_count,
                     const int64_t *const *keys, const int *key_moduli_ids,
                     const int *deltas, int64_t *const *outs) {
  return guarded([&] {
    require(ctx && in && keys && key_moduli_ids && deltas && outs,
            "null argument");
    require(count > 0 && key_count > 0, "Rotate: empty batch");
    ctx->check_moduli(moduli_id, 1, "Rotate");
    for (int k = 0; k < key_count; k++) {
      require(keys[k] && outs[k], "null argument");
      require(static_cast<const void *>(outs[k]) !=
                  static_cast<const void *>(in),
              "Rotate: a batch output must not alias the input");
      ctx->check_key(moduli_id, key_moduli_ids[k], "Rotate");
    }
    const auto &P = ctx->params;
    const int limbs = P.limbs_at(moduli_id);
    const int ext_limbs = limbs + P.alpha;
    const int64_t N = ctx->n();
    const int64_t duo_l = int64_t(count) * 2 * limbs;
    const int64_t ext_l =
        int64_t(count) * ctx->ext_slots(moduli_id) * ext_limbs;
    const int64_t acc_l = int64_t(count) * 2 * ext_limbs;
    int64_t *ws = ctx->scratch(size_t(duo_l + ext_l + acc_l) * N * 8);
    int64_t *sw = ws, *ext = sw + duo_l * N, *acc = ext + ext_l * N;
    const bool hoist = fhe::ks_fused_supported(ctx->tables);
    // CKKSEngine::rotate_batch (src/ckks/engine_rotate.cpp:262-300): one
    // decomposition of c1, then the key-dependent part per key
    if (hoist)
      ctx->ks_decompose(count, moduli_id, key_moduli_ids[0], in + limbs * N,
                        2 * limbs * N, ext);
    for (int k = 0; k < key_count; k++) {
      if (hoist)
        ctx->ks_apply(count, moduli_id, keys[k], key_moduli_ids[k], ext, acc);
      else
        ctx->key_switch(count, moduli_id, keys[k], key_moduli_ids[k],
                        in + limbs * N, 2 * limbs * N, ext, acc);
      ctx->stage("moddown_add", [&] {
        fhe::launch_moddown_add(ctx->tables, acc, in, 2 * limbs * N, 1, sw,
                                count, limbs, moduli_id, ctx->stream);
      });
      ctx->stage("automorphism", [&] {
        fhe::launch_automorphism(ctx->tables, sw, outs[k], duo_l, limbs,
                                 moduli_id,
                                 automorphism_exponent(P.log_n, deltas[k]),
                                 ctx->stream);
      });
    }
    ctx->check_launch("Rotate");
  });
}

int fhe_switch_key(fhe_context *ctx, int count, int moduli_id,
                   const int64_t *in, const int64_t *key, int key_moduli_id,
                   int64_t *out) {
  return guarded([&] {
    require(ctx && in && key && out, "null argument");
    require(count > 0, "Switch key: empty batch");
    ctx->check_moduli(moduli_id, 1, "Switch key");
    ctx->check_key(moduli_id, key_moduli_id, "Switch key");
    ctx->switch_key_into(count, moduli_id, in, key, key_moduli_id, out);
    ctx->check_launch("Switch key");
  });
}

int fhe_multiply_imaginary(fhe_context *ctx, int count, int moduli_id,
                           const int64_t *in, int64_t value, int64_t *out) {
  return guarded([&] {
    require(ctx && in && out, "null argument");
    require(count > 0, "Multiply imaginary: empty batch");
    ctx->check_moduli(moduli_id, 1, "Multiply imaginary");
    const auto &P = ctx->params;
    const int limbs = P.limbs_at(moduli_id);
    const int64_t N = ctx->n();
    const int64_t poly = int64_t(limbs) * N;
    // the plaintext value * X^(N/2) (CKKSCore::encode_imaginary,
    // src/ckks/core_encode.cpp:190-210), tiled on the device and transformed
    int64_t *pt = ctx->stash(size_t(poly) * 8);
    ctx->stage("multiply_imaginary", [&] {
      fhe::launch_tile_monomial(ctx->tables, value, N / 2, pt, limbs,
                                moduli_id, ctx->stream);
      fhe::ntt_forward(ctx->tables, pt, pt, limbs, limbs, moduli_id,
                       ctx->tables.r2, ctx->stream);
      // ciphertext to NTT form, product, back to INTT form
      // (src/ckks/engine_multiply.cpp:468-491)
      fhe::ntt_forward(ctx->tables, in, out, int64_t(count) * 2 * limbs, limbs,
                       moduli_id, ctx->tables.r2, ctx->stream);
      for (int c = 0; c < 2 * count; c++)
        fhe::launch_mont_mult(ctx->tables, out + c * poly, pt, out + c * poly,
                              limbs, limbs, moduli_id, ctx->stream);
      fhe::ntt_inverse(ctx->tables, out, out, int64_t(count) * 2 * limbs,
                       limbs, moduli_id, fhe::kInttExit, ctx->stream);
    });
    ctx->check_launch("Multiply imaginary");
  });
}

int fhe_mod_raise(fhe_context *ctx, int count, const int64_t *in,
                  const int64_t *to_sparse, int sparse_moduli_id,
                  const int64_t *to_dense, int dense_moduli_id,
                  int64_t *out) {
  return guarded([&] {
    require(ctx && in && to_sparse && to_dense && out, "null argument");
    require(count > 0, "Mod raise: empty batch");
    const auto &P = ctx->params;
    // input: level 0, not rescaled (moduli_id full_level, two limbs), the
    // form CKKSEngine::mod_raise hands over after level_down(ct, 0); the
    // rescale leaves the base prime alone
    const int low = P.full_level + 1;
    const int high = P.moduli_id_of_level(P.max_key_switching_level());
    require(low < P.chain_count, "Mod raise: no level below level 0");
    ctx->check_key(low, sparse_moduli_id, "Mod raise");
    ctx->check_key(high, dense_moduli_id, "Mod raise");
    const int64_t N = ctx->n();
    const int high_limbs = P.limbs_at(high);
    // stash: the rescaled duo, the switched duo, its signed limbs (one limb
    // each) and the tiled duo
    int64_t *r1 = ctx->stash(size_t(count) * (2 + 2 + 2 + 2 * high_limbs) *
                             N * 8);
    int64_t *sw = r1 + 2 * int64_t(count) * N;
    int64_t *sg = sw + 2 * int64_t(count) * N;
    int64_t *tiled = sg + 2 * int64_t(count) * N;
    // CKKSEngine::mod_raise_ (src/ckks/engine_mod_raise.cpp:5-34)
    ctx->stage("rescale", [&] {
      fhe::launch_rescale(ctx->tables, in, r1, 2 * count, 2, P.full_level,
                          ctx->stream);
    });
    ctx->switch_key_into(count, low, r1, to_sparse, sparse_moduli_id, sw);
    // CKKSCore::make_signed_tile (src/ckks/core.cpp:1085-1110)
    ctx->stage("elementwise", [&] {
      fhe::launch_elementwise(ctx->tables, fhe::kEwMakeSigned, sw, nullptr, sg,
                              2 * int64_t(count), 1, low, nullptr,
                              ctx->stream);
      fhe::launch_tile_unsigned(ctx->tables, sg, tiled, 2 * count, high_limbs,
                                high, ctx->stream);
    });
    // the tiled limbs are signed values, not residues
    ctx->switch_key_into(count, high, tiled, to_dense, dense_moduli_id, out,
                         true);
    // multiply by the integer ratio (CKKSCore::get_mod_raise_ratio,
    // src/ckks/core.cpp:1112-1119)
    const auto c = P.mult_multiplier_int(P.mod_raise_ratio(), high);
    require(int(c.size()) <= fhe::kMaxEwLimbs, "Mod raise: too many limbs");
    ctx->stage("elementwise", [&] {
      fhe::launch_elementwise(ctx->tables, fhe::kEwEnterReduce, out, nullptr,
                              out, 2 * int64_t(count) * high_limbs, high_limbs,
                              high, c.data(), ctx->stream);
    });
    ctx->check_launch("Mod raise");
  });
}

int fhe_conjugate(fhe_context *ctx, int count, int moduli_id,
                  const int64_t *in, const int64_t *key, int key_moduli_id,
                  int64_t *out) {
  return guarded([&] {
    require(ctx && in && key && out, "null argument");
    require(count > 0, "Conjugate: empty batch");
    ctx->check_moduli(moduli_id, 1, "Conjugate");
    ctx->check_key(moduli_id, key_moduli_id, "Conjugate");
    const auto &P = ctx->params;
    const int limbs = P.limbs_at(moduli_id);
    const int ext_limbs = limbs + P.alpha;
    const int64_t N = ctx->n();
    const int64_t duo_l = int64_t(count) * 2 * limbs;
    const int64_t ext_l =
        int64_t(count) * ctx->ext_slots(moduli_id) * ext_limbs;
    const int64_t acc_l = int64_t(count) * 2 * ext_limbs;
    int64_t *ws = ctx->scratch(size_t(duo_l + ext_l + acc_l) * N * 8);
    int64_t *perm = ws, *ext = perm + duo_l * N, *acc = ext + ext_l * N;
    // CKKSCore::conjugate (src/ckks/core.cpp:1035-1056): automorphism
    // first, then key switching of the permuted c1 plus the permuted c0
    ctx->stage("automorphism", [&] {
      fhe::launch_automorphism(ctx->tables, in, perm, duo_l, limbs, moduli_id,
                               automorphism_exponent(P.log_n, -1),
                               ctx->stream);
    });
    ctx->key_switch(count, moduli_id, key, key_moduli_id, perm + limbs * N,
                    2 * limbs * N, ext, acc);
    ctx->stage("moddown_add", [&] {
      fhe::launch_moddown_add(ctx->tables, acc, perm, 2 * limbs * N, 1, out,
                              count, limbs, moduli_id, ctx->stream);
    });
    ctx->check_launch("Conjugate");
  });
}

} // extern "C"

