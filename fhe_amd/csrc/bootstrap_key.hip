// Bootstrapping with a BootstrapKey on the device: the key (src/data/
// bootstrap.h:24-44) holds, per butterfly stage, a PlainMatrix
// (src/data/plain_matrix.h:10-24: the merged diagonals encoded once in the
// extended basis, rescaled and in NTT form, CKKSEngine::make_bootstrap_stage,
// src/ckks/engine_bootstrap.cpp:8-69) and one fixed rotation key per delta.
// A stage (apply_butterfly, src/ckks/engine_bootstrap.cpp:218-289) then
// decomposes its input once for all baby steps and keeps every rotation in
// the extended basis until a single ModDown at the end:
//
//   rescale; decompose c1 (ModUp, NTT); attach the special primes to the
//   input (x P, NTT); baby steps = ntt_rotate of the attached input with the
//   shared decomposition and each inner delta's key; per giant-step group
//   (equal outer deltas), the sum of the baby steps' products with the
//   group's diagonals; each group with a nonzero delta is ModDown'ed,
//   decomposed again and ntt_rotate'd by its delta; the groups' sum leaves
//   the extended basis by INTT and ModDown.
//
// Device layout of one stage (B ciphertexts, E = limbs + alpha extended
// limbs, parts partitions):
//   attached  [B][2][E][N]            decomposition [B][parts][E][N]
//   baby steps [B][inner - 1][2][E][N]
//   group sums [2][groups][B][E][N]   groups with a nonzero delta first
//   giant decompositions [groups'][B][parts][E][N]
//   outer sum [B][2][E][N] -> result [B][2][limbs][N]
// Everything in the extended basis is read only by ModDown after an INTT
// reduced to [0, q), so the kernels keep any representative below 2q.
#include <map>

#include "text_engine.h"

namespace {

struct BkStage {
  int level = 0;
  int inner = 0;
  int outer = 0;
  std::vector<int> inner_deltas, outer_deltas;
  // giant-step groups (PlainMatrix::outer_loop_process_order_): the distinct
  // outer deltas in the order of first appearance, each with its outer
  // indices
  std::vector<std::pair<int, std::vector<int>>> order;
  const int64_t **table = nullptr; // device [outer][inner] diagonal pointers
};

struct RotationEntry {
  const int64_t *key = nullptr;
  int moduli_id = 0;
};

} // namespace

struct fhe_bootstrap_key {
  fhe_context *ctx = nullptr;
  int stage_count = 0;
  std::vector<BkStage> cts, stc;
  std::map<int, RotationEntry> rotations;
  const int64_t *to_sparse = nullptr;
  int to_sparse_moduli_id = 0;
  const int64_t *to_dense = nullptr;
  int to_dense_moduli_id = 0;
  // P R^2 mod q per prime (P the special primes' product): the input
  // multiplier of the forward NTT that attaches the special primes
  int64_t *p_r2 = nullptr;

  ~fhe_bootstrap_key() {
    if (ctx)
      (void)hipStreamSynchronize(ctx->stream);
    for (auto *list : {&cts, &stc})
      for (auto &st : *list)
        if (st.table)
          (void)hipFree(st.table);
    if (p_r2)
      (void)hipFree(p_r2);
  }
};

namespace {

using U128 = unsigned __int128;

// get_stage_loop_counts / make_bootstrap_stage's grouping of outer indices
// by delta (src/ckks/engine_bootstrap.cpp:34-47)
std::vector<std::pair<int, std::vector<int>>>
process_order(const std::vector<int> &outer_deltas) {
  std::vector<std::pair<int, std::vector<int>>> order;
  for (int o = 0; o < int(outer_deltas.size()); o++) {
    bool found = false;
    for (auto &entry : order)
      if (entry.first == outer_deltas[o]) {
        entry.second.push_back(o);
        found = true;
        break;
      }
    if (!found)
      order.push_back({outer_deltas[o], {o}});
  }
  return order;
}

BkStage make_stage(fhe_context *ctx, const fhe_plain_matrix &m,
                   const fhe_bootstrap_key &bk, const char *what) {
  const auto &P = ctx->params;
  require(m.inner_count >= 1 && m.outer_count >= 1 &&
              m.outer_count <= fhe::kBkMaxOuter &&
              int64_t(m.inner_count) * m.outer_count <= 4096,
          std::string("Bootstrap key: the loop counts of a ") + what +
              " stage are out of range");
  require(m.inner_deltas && m.outer_deltas && m.diagonals,
          "Bootstrap key: a stage's deltas or diagonals are missing");
  // the stage's diagonals sit at its level, rescaled
  require(m.level >= 1 && m.level <= P.full_level &&
              P.moduli_id_of_level(m.level) + 1 < P.chain_count,
          "Bootstrap key: a stage's level is out of range");
  BkStage st;
  st.level = m.level;
  st.inner = m.inner_count;
  st.outer = m.outer_count;
  st.inner_deltas.assign(m.inner_deltas, m.inner_deltas + m.inner_count);
  st.outer_deltas.assign(m.outer_deltas, m.outer_deltas + m.outer_count);
  require(st.inner_deltas[0] == 0,
          "Bootstrap key: a stage's first inner delta is not 0");
  for (int i = 1; i < st.inner; i++)
    require(bk.rotations.count(st.inner_deltas[i]) != 0,
            "Bootstrap key: no rotation key for inner delta " +
                std::to_string(st.inner_deltas[i]));
  st.order = process_order(st.outer_deltas);
  for (const auto &g : st.order)
    require(g.first == 0 || bk.rotations.count(g.first) != 0,
            "Bootstrap key: no rotation key for outer delta " +
                std::to_string(g.first));
  const int mid = P.moduli_id_of_level(m.level) + 1;
  for (int i = 1; i < st.inner; i++)
    require(bk.rotations.at(st.inner_deltas[i]).moduli_id <= mid,
            "Bootstrap key: a rotation key sits below its stage's level");
  std::vector<const int64_t *> host(size_t(st.outer) * st.inner, nullptr);
  for (int o = 0; o < st.outer; o++)
    for (int i = (o == 0 ? 1 : 0); i < st.inner; i++) {
      const int64_t *d = m.diagonals[o * st.inner + i];
      require(d != nullptr, "Bootstrap key: diagonal (" + std::to_string(o) +
                                ", " + std::to_string(i) + ") is missing");
      host[size_t(o) * st.inner + i] = d;
    }
  void *p = nullptr;
  hip_check(hipMalloc(&p, host.size() * sizeof(void *)),
            "bootstrap key allocation");
  st.table = static_cast<const int64_t **>(p);
  hip_check(hipMemcpy(p, host.data(), host.size() * sizeof(void *),
                      hipMemcpyHostToDevice),
            "bootstrap key upload");
  return st;
}

// One ntt_rotate term by `delta` of a decomposition at moduli_id m.
fhe::BkTerm rotation_term(const fhe_context *ctx, const fhe_bootstrap_key &bk,
                          int m, int delta, const int64_t *c0,
                          int64_t src_stride, const int64_t *dec,
                          int64_t dec_stride) {
  const auto &P = ctx->params;
  const auto it = bk.rotations.find(delta);
  require(it != bk.rotations.end(),
          "Bootstrap: the bootstrap key does not contain a rotation context "
          "for delta " +
              std::to_string(delta) + ".");
  const int kmid = it->second.moduli_id;
  require(kmid >= 0 && kmid <= m,
          "Bootstrap: the key level is lower than the ciphertext level.");
  const int key_chain = P.limbs_at(kmid) + P.alpha;
  const int key_offset = P.limbs_at(kmid) - P.limbs_at(m);
  const auto parts = P.partitions(m, kmid);
  fhe::BkTerm t{};
  t.c0 = c0;
  t.dec = dec;
  t.key = it->second.key +
          (int64_t(2 * parts[0].key_index) * key_chain + key_offset) * P.n;
  t.src_stride = src_stride;
  t.dec_stride = dec_stride;
  t.key_chain = key_chain;
  t.g = automorphism_exponent(P.log_n, delta);
  return t;
}

// ks_decompose and the second forward pass: the full NTT-form decomposition
// (CKKSCore::decompose, src/ckks/core_key_switch.cpp:434-460) of `count`
// chain polynomials at moduli_id m.
void decompose_full(fhe_context *ctx, int count, int m, const int64_t *src,
                    int64_t src_stride, int64_t *dec) {
  const auto &P = ctx->params;
  const int ext_limbs = P.limbs_at(m) + P.alpha;
  const int parts = int(P.partitions(m, m).size());
  ctx->ks_decompose(count, m, m, src, src_stride, dec);
  const fhe::LazySpec lazy =
      ctx->lazy_key_switch ? fhe::kAllLazy : fhe::LazySpec();
  ctx->stage("bk_ntt_pass_low(ext)", [&] {
    fhe::ntt_forward_second(ctx->tables, dec, dec,
                            int64_t(count) * parts * ext_limbs, ext_limbs, m,
                            ctx->stream, lazy);
  });
}

// apply_butterfly with a stage (src/ckks/engine_bootstrap.cpp:218-289) of
// every ciphertext of the batch.
Text apply_butterfly_bk(TextEngine &e, const fhe_bootstrap_key &bk,
                        const Text &x, const BkStage &st) {
  fhe_context *ctx = e.context();
  const auto &P = ctx->params;
  const auto &T = ctx->tables;
  hipStream_t stream = ctx->stream;
  require(x.polys == 2 && !(x.flags & FHE_FLAG_NTT_FORM),
          "Bootstrap: a coefficient-domain duo enters a butterfly stage");
  const Text r = e.rescale(x);
  const int m = ctx->moduli_id_of(r.level, true);
  require(m == P.moduli_id_of_level(st.level) + 1,
          "Bootstrap: the ciphertext's level differs from the stage's");
  require(m + 1 < P.chain_count, "Bootstrap: the level is too low.");
  const int B = r.count;
  const int limbs = P.limbs_at(m), E = limbs + P.alpha;
  const int64_t N = ctx->n(), ext = int64_t(E) * N;
  const int parts = int(P.partitions(m, m).size());

  // attach_special_prime_chains_ (src/ckks/core.cpp:1127-1155): x P in the
  // chain limbs, zero special limbs, NTT
  auto attached_h = e.take(size_t(B) * 2 * ext * 8);
  int64_t *attached = static_cast<int64_t *>(attached_h.get());
  ctx->stage("bk_attach", [&] {
    fhe::launch_bk_attach(T, r.data, attached, int64_t(B) * 2, limbs, E,
                          stream);
    fhe::ntt_forward(T, attached, attached, int64_t(B) * 2 * E, E, m,
                     bk.p_r2, stream);
  });
  auto dec_h = e.take(size_t(B) * parts * ext * 8);
  int64_t *dec = static_cast<int64_t *>(dec_h.get());
  decompose_full(ctx, B, m, r.data + int64_t(limbs) * N,
                 int64_t(2) * limbs * N, dec);

  // baby steps: ntt_rotate of the attached input by every inner delta
  const int babies = st.inner - 1;
  std::shared_ptr<void> rot_h;
  int64_t *rot = nullptr;
  if (babies > 0) {
    rot_h = e.take(size_t(B) * babies * 2 * ext * 8);
    rot = static_cast<int64_t *>(rot_h.get());
  }
  for (int i0 = 1; i0 < st.inner; i0 += fhe::kBkMaxTerms) {
    fhe::BkRotateArgs a{};
    a.count = B;
    a.parts = parts;
    a.ext_limbs = E;
    a.moduli_id = m;
    a.out = rot + int64_t(i0 - 1) * 2 * ext;
    a.out_c_stride = int64_t(babies) * 2 * ext;
    a.out_t_stride = 2 * ext;
    for (int i = i0; i < std::min(st.inner, i0 + fhe::kBkMaxTerms); i++)
      a.term[a.terms++] =
          rotation_term(ctx, bk, m, st.inner_deltas[i], attached, 2 * ext, dec,
                        int64_t(parts) * ext);
    ctx->stage("bk_baby_steps", [&] { fhe::launch_bk_rotate(T, a, false, stream); });
  }

  // products with the diagonals, summed per group; groups with a nonzero
  // delta first
  std::vector<int> slot_group;
  for (int g = 0; g < int(st.order.size()); g++)
    if (st.order[g].first != 0)
      slot_group.push_back(g);
  const int rotating = int(slot_group.size());
  for (int g = 0; g < int(st.order.size()); g++)
    if (st.order[g].first == 0)
      slot_group.push_back(g);
  const int G = int(slot_group.size());
  auto sums_h = e.take(size_t(2) * G * B * ext * 8);
  int64_t *sums = static_cast<int64_t *>(sums_h.get());
  for (int s0 = 0; s0 < G; s0 += fhe::kBkMaxGroups) {
    fhe::BkDiagArgs d{};
    d.base = attached;
    d.rotated = rot;
    d.diagonal = st.table;
    d.sums = sums;
    d.inner = st.inner;
    d.count = B;
    d.ext_limbs = E;
    d.moduli_id = m;
    d.groups = G;
    d.group0 = s0;
    d.groups_here = std::min(fhe::kBkMaxGroups, G - s0);
    int members = 0;
    for (int g = 0; g < d.groups_here; g++) {
      d.member_start[g] = members;
      for (int o : st.order[slot_group[s0 + g]].second)
        d.member[members++] = o;
    }
    d.member_start[d.groups_here] = members;
    ctx->stage("bk_diagonals", [&] { fhe::launch_bk_diagonal(T, d, stream); });
  }

  // giant steps: mod_down_then_decompose (src/ckks/core_key_switch.cpp:
  // 467-505) of every rotating group's polynomial 1, then ntt_rotate
  std::shared_ptr<void> decg_h, reduced_h;
  int64_t *decg = nullptr;
  if (rotating > 0) {
    int64_t *c1s = sums + int64_t(G) * B * ext;
    reduced_h = e.take(size_t(rotating) * B * limbs * N * 8);
    int64_t *reduced = static_cast<int64_t *>(reduced_h.get());
    ctx->stage("bk_giant_moddown", [&] {
      fhe::ntt_inverse(T, c1s, c1s, int64_t(rotating) * B * E, E, m,
                       fhe::kInttExitReduce, stream);
      fhe::launch_moddown_polys(T, c1s, reduced, int64_t(rotating) * B, limbs,
                                m, stream);
    });
    decg_h = e.take(size_t(rotating) * B * parts * ext * 8);
    decg = static_cast<int64_t *>(decg_h.get());
    decompose_full(ctx, rotating * B, m, reduced, int64_t(limbs) * N, decg);
  }
  auto outer_h = e.take(size_t(B) * 2 * ext * 8);
  int64_t *outer = static_cast<int64_t *>(outer_h.get());
  for (int k0 = 0; k0 < G; k0 += fhe::kBkMaxTerms) {
    fhe::BkRotateArgs a{};
    a.count = B;
    a.parts = parts;
    a.ext_limbs = E;
    a.moduli_id = m;
    a.accumulate = k0 > 0;
    a.out = outer;
    a.out_c_stride = 2 * ext;
    for (int k = k0; k < std::min(G, k0 + fhe::kBkMaxTerms); k++) {
      const int64_t *c0 = sums + int64_t(k) * B * ext;
      if (k < rotating) {
        a.term[a.terms++] = rotation_term(
            ctx, bk, m, st.order[slot_group[k]].first, c0, ext,
            decg + int64_t(k) * B * parts * ext, int64_t(parts) * ext);
      } else {
        fhe::BkTerm t{};
        t.c0 = c0;
        t.c1 = sums + (int64_t(G) + k) * B * ext;
        t.src_stride = ext;
        t.g = 1;
        a.term[a.terms++] = t;
      }
    }
    ctx->stage("bk_giant_steps", [&] { fhe::launch_bk_rotate(T, a, true, stream); });
  }

  // remove_special_prime_chains_ (src/ckks/core.cpp:1165-1187)
  const int out_level = r.level - 1;
  require(ctx->moduli_id_of(out_level, false) == m,
          "Bootstrap: unexpected level bookkeeping");
  Text out = e.alloc(2, out_level, 0, B);
  ctx->stage("bk_remove_special", [&] {
    fhe::ntt_inverse(T, outer, outer, int64_t(B) * 2 * E, E, m,
                     fhe::kInttExitReduce, stream);
    fhe::launch_moddown_add(T, outer, nullptr, 0, 0, out.data, B, limbs, m,
                            stream);
  });
  return out;
}

// coeff_to_slot with a BootstrapKey (src/ckks/engine_bootstrap.cpp:368-398)
std::pair<Text, Text> coeff_to_slot_bk(TextEngine &e,
                                       const fhe_bootstrap_key &bk,
                                       const Text &x) {
  require(!bk.cts.empty(), "Coeff to slot: the bootstrap key has no stages");
  Text current = x;
  for (const BkStage &st : bk.cts)
    current = apply_butterfly_bk(e, bk, current, st);
  const Text conjugated = e.conjugate(current);
  Text real = e.add(current, conjugated, false);
  const Text subtracted = e.add(current, conjugated, true);
  Text imag = e.multiply_imaginary(subtracted, -1);
  return {real, imag};
}

// slot_to_coeff with a BootstrapKey (src/ckks/engine_bootstrap.cpp:436-454)
Text slot_to_coeff_bk(TextEngine &e, const fhe_bootstrap_key &bk,
                      const Text &real, const Text &imag) {
  require(!bk.stc.empty(), "Slot to coeff: the bootstrap key has no stages");
  const Text multiplied = e.multiply_imaginary(imag, 1);
  Text current = e.add(real, multiplied, false);
  for (const BkStage &st : bk.stc)
    current = apply_butterfly_bk(e, bk, current, st);
  return current;
}

// rotate(ct, BootstrapKey, delta, false) (src/ckks/engine_rotate.cpp:
// 190-205): the fixed key of the delta normalized by N / 2
Text rotate_bk(TextEngine &e, const fhe_bootstrap_key &bk, const Text &a,
               int delta) {
  fhe_context *ctx = e.context();
  const int d = e.normalized(delta, false);
  if (d == 0)
    return e.clone(a);
  const auto it = bk.rotations.find(d);
  require(it != bk.rotations.end(),
          "Rotate: the bootstrap key does not contain a rotation context for "
          "delta " +
              std::to_string(d) + ".");
  const Text in = e.key_switch_input(a);
  const int mid = ctx->moduli_id_of(in.level, in.flags & FHE_FLAG_RESCALED);
  Text out = e.alloc(in.polys, in.level, in.flags, in.count);
  ok(fhe_rotate(ctx, in.count, mid, in.data, it->second.key,
                it->second.moduli_id, d, out.data));
  return out;
}

// The keys of a bootstrap with a BootstrapKey as the TextEngine reads them.
fhe_bootstrap_keys engine_keys(const fhe_bootstrap_key &bk,
                               const int64_t *relin, int relin_moduli_id,
                               const int64_t *conj, int conj_moduli_id) {
  fhe_bootstrap_keys k{};
  k.relin = relin;
  k.relin_moduli_id = relin_moduli_id;
  k.conjugation = conj;
  k.conjugation_moduli_id = conj_moduli_id;
  k.to_sparse = bk.to_sparse;
  k.to_sparse_moduli_id = bk.to_sparse_moduli_id;
  k.to_dense = bk.to_dense;
  k.to_dense_moduli_id = bk.to_dense_moduli_id;
  return k;
}

void require_bk(const fhe_context *ctx, const fhe_bootstrap_key *bk) {
  require(bk != nullptr, "Bootstrap: the bootstrap key is missing");
  require(bk->ctx == ctx, "Bootstrap: the bootstrap key belongs to another "
                          "context");
}

} // namespace

extern "C" {

int fhe_bootstrap_key_create(fhe_context *ctx,
                             const fhe_bootstrap_key_parts *parts,
                             fhe_bootstrap_key **out) {
  return guarded([&] {
    require(ctx && parts && out, "null argument");
    *out = nullptr;
    ctx->require_not_capturing("Bootstrap key");
    const auto &P = ctx->params;
    require(fhe::bk_supported(ctx->tables),
            "Bootstrap key: N is too small");
    require(parts->stage_count >= 1, "Bootstrap key: stage count");
    require(parts->rotation_count >= 0 &&
                (parts->rotation_count == 0 ||
                 (parts->rotation_deltas && parts->rotation_keys &&
                  parts->rotation_moduli_ids)),
            "Bootstrap key: rotation keys missing");
    auto bk = std::make_unique<fhe_bootstrap_key>();
    bk->ctx = ctx;
    bk->stage_count = parts->stage_count;
    for (int i = 0; i < parts->rotation_count; i++) {
      const int mid = parts->rotation_moduli_ids[i];
      require(parts->rotation_keys[i] != nullptr,
              "Bootstrap key: rotation key missing");
      require(mid >= 0 && mid < P.chain_count &&
                  P.max_key_switching_level() >= P.level_of(mid, false),
              "Bootstrap key: a rotation key's moduli id is out of range");
      require(bk->rotations.count(parts->rotation_deltas[i]) == 0,
              "Bootstrap key: a delta appears twice");
      bk->rotations[parts->rotation_deltas[i]] = {parts->rotation_keys[i],
                                                  mid};
    }
    bk->to_sparse = parts->to_sparse;
    bk->to_sparse_moduli_id = parts->to_sparse_moduli_id;
    bk->to_dense = parts->to_dense;
    bk->to_dense_moduli_id = parts->to_dense_moduli_id;
    for (int s = 0; s < parts->stage_count; s++) {
      if (parts->coeff_to_slot)
        bk->cts.push_back(
            make_stage(ctx, parts->coeff_to_slot[s], *bk, "coeff-to-slot"));
      if (parts->slot_to_coeff)
        bk->stc.push_back(
            make_stage(ctx, parts->slot_to_coeff[s], *bk, "slot-to-coeff"));
    }
    // P R^2 mod q for every prime (0 for the special primes themselves)
    std::vector<int64_t> pr2(size_t(P.prime_count));
    for (int p = 0; p < P.prime_count; p++) {
      const uint64_t q = P.q[size_t(p)];
      uint64_t prod = 1;
      for (int j = P.chain_count; j < P.prime_count; j++)
        prod = uint64_t(U128(prod) * (P.q[size_t(j)] % q) % q);
      const uint64_t r = uint64_t((U128(1) << 62) % q);
      const uint64_t r2 = uint64_t(U128(r) * r % q);
      pr2[size_t(p)] = int64_t(U128(prod) * r2 % q);
    }
    void *p = nullptr;
    hip_check(hipMalloc(&p, pr2.size() * 8), "bootstrap key allocation");
    bk->p_r2 = static_cast<int64_t *>(p);
    hip_check(hipMemcpy(p, pr2.data(), pr2.size() * 8, hipMemcpyHostToDevice),
              "bootstrap key upload");
    // the decompositions at every stage level use their own level's ModUp
    // constants: uploaded now, outside any graph capture
    for (const auto *list : {&bk->cts, &bk->stc})
      for (const BkStage &st : *list) {
        const int m = P.moduli_id_of_level(st.level) + 1;
        (void)ctx->modup_constants(m, m);
      }
    *out = bk.release();
  });
}

void fhe_bootstrap_key_destroy(fhe_bootstrap_key *bk) { delete bk; }

int fhe_coeff_to_slot_bk(fhe_context *ctx, int count, const int64_t *in,
                         int level, int flags, const int64_t *conj,
                         int conj_moduli_id, const fhe_bootstrap_key *bk,
                         int64_t *real, int64_t *imag, int *out_level,
                         int *out_flags) {
  return guarded([&] {
    require(ctx && in && conj && real && imag && out_level && out_flags,
            "null argument");
    require_bk(ctx, bk);
    require(count > 0, "Bootstrap: empty batch");
    require_boot_form(ctx->params, level, flags, 1, "Coeff to slot");
    const fhe_bootstrap_keys keys =
        engine_keys(*bk, nullptr, 0, conj, conj_moduli_id);
    TextEngine e(ctx, nullptr, 0);
    e.keys = &keys;
    e.count = count;
    const auto [r, i] = coeff_to_slot_bk(e, *bk, e.wrap(in, 2, level, flags));
    int l2, f2;
    copy_out(ctx, e, r, real, out_level, out_flags);
    copy_out(ctx, e, i, imag, &l2, &f2);
    require(l2 == *out_level && f2 == *out_flags, "Coeff to slot: forms");
    ctx->check_launch("Coeff to slot");
  });
}

int fhe_slot_to_coeff_bk(fhe_context *ctx, int count, const int64_t *real,
                         const int64_t *imag, int level, int flags,
                         const int64_t *conj, int conj_moduli_id,
                         const fhe_bootstrap_key *bk, int64_t *out,
                         int *out_level, int *out_flags) {
  return guarded([&] {
    require(ctx && real && imag && out && out_level && out_flags,
            "null argument");
    require_bk(ctx, bk);
    require(count > 0, "Bootstrap: empty batch");
    require_boot_form(ctx->params, level, flags, 1, "Slot to coeff");
    const fhe_bootstrap_keys keys =
        engine_keys(*bk, nullptr, 0, conj, conj_moduli_id);
    TextEngine e(ctx, nullptr, 0);
    e.keys = &keys;
    e.count = count;
    const Text c = slot_to_coeff_bk(e, *bk, e.wrap(real, 2, level, flags),
                                    e.wrap(imag, 2, level, flags));
    copy_out(ctx, e, c, out, out_level, out_flags);
    ctx->check_launch("Slot to coeff");
  });
}

int fhe_bootstrap_bk(fhe_context *ctx, int count, const int64_t *in, int level,
                     int flags, const int64_t *relin, int relin_moduli_id,
                     const int64_t *conj, int conj_moduli_id,
                     const fhe_bootstrap_key *bk, int64_t *out, int *out_level,
                     int *out_flags) {
  return guarded([&] {
    require(ctx && in && relin && conj && out && out_level && out_flags,
            "null argument");
    require_bk(ctx, bk);
    require(bk->to_sparse && bk->to_dense,
            "Bootstrap: the small bootstrap key is missing");
    require(!bk->cts.empty() && !bk->stc.empty(),
            "Bootstrap: the bootstrap key lacks stages");
    require(!(flags & FHE_FLAG_NTT_FORM), "Bootstrap: an INTT-form duo");
    require(count > 0, "Bootstrap: empty batch");
    require_boot_form(ctx->params, level, flags, 0, "Bootstrap");
    const fhe_bootstrap_keys keys =
        engine_keys(*bk, relin, relin_moduli_id, conj, conj_moduli_id);
    TextEngine e(ctx, relin, relin_moduli_id);
    e.keys = &keys;
    e.count = count;
    const int slots = int(ctx->params.slot_count());
    BootstrapSteps steps;
    // partial_sum_inplace with the bootstrap key
    // (src/ckks/engine_bootstrap.cpp:291-313)
    steps.partial_sum = [&](Text &raised) {
      for (int delta = slots; delta < ctx->n() / 2; delta <<= 1) {
        const Text r = rotate_bk(e, *bk, raised, delta);
        e.add_inplace(raised, r);
      }
    };
    steps.coeff_to_slot = [&](const Text &raised) {
      return coeff_to_slot_bk(e, *bk, raised);
    };
    steps.slot_to_coeff = [&](const Text &r, const Text &i) {
      return slot_to_coeff_bk(e, *bk, r, i);
    };
    const Text result = bootstrap_texts(e, e.wrap(in, 2, level, flags), steps);
    copy_out(ctx, e, result, out, out_level, out_flags);
    ctx->check_launch("Bootstrap");
  });
}

} // extern "C"
