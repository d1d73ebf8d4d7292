// Key switching of batches at uneven levels: the C ABI, the per-shape work
// lists (plans), their autotuned launch schedules, and HIP graph capture of
// a context's stream.
//
// The reference switches keys one ciphertext at a time
// (CKKSCore::relinearize, src/ckks/core.cpp:713-741; CKKSCore::rotate,
// src/ckks/core.cpp:938-962; CKKSEngine::rotate with a RotationKey,
// src/ckks/engine_rotate.cpp:142-170), each through
// create_key_switcher (src/ckks/core_key_switch.cpp:293-358) and mod_down
// (:110-160).  Here a batch of ciphertexts at different moduli_ids goes
// through the same stages with one launch per stage for the whole batch
// (uneven.h): each ciphertext's limbs get the arithmetic they get in the
// batched kernels of one moduli_id, so every output limb equals the
// ciphertext's own single call.  Blocks that work on the same absolute prime
// and tile run next to each other whatever their ciphertext's level, so the
// key tile (key limb = absolute prime - key moduli_id for every level) and
// the prime's twiddles are read from HBM once per batch.
#include "context.h"
#include "ks_tile.h"
#include "uneven.h"

namespace {

using fhe::UnevenCt;
using fhe::UnevenPtrs;
using fhe::UnitList;

// Pipeline stages of one uneven key switch.
enum UnevenStage : int {
  kStTrioIntt1 = 0, // relinearize: inverse pass 1 of the trio
  kStTrioIntt2,     //              inverse pass 2 with exit
  kStDigits,        // Garner digits of every partition
  kStExtend,        // extended limbs + forward pass 1
  kStKs,            // forward pass 2, key products, inverse pass 1
  kStAccIntt2,      // inverse pass 2 of both accumulators
  kStModDown,       // ModDown + addend
  kStAutomorphism,  // rotation: X -> X^g
  kStCount
};

const char *const kStageNames[kStCount] = {
    "uneven_intt1(trio)", "uneven_intt2(trio)", "uneven_digits",
    "uneven_extend",      "uneven_ks",          "uneven_intt2(acc)",
    "uneven_moddown",     "uneven_automorphism"};

// Launch schedule candidates: persistent grid caps (0: one block per unit)
// with and without the per-XCD ranges.
struct Schedule {
  int cap, xcd;
};
const Schedule kCandidates[] = {{0, 1}, {0, 0}, {1024, 1}, {2048, 1},
                                {4096, 1}, {2048, 0}};
constexpr int kCandidateCount = int(sizeof(kCandidates) / sizeof(Schedule));

} // namespace

// The work lists of one batch shape.
struct UnevenPlan {
  bool relin = false; // NTT-form trios (own limbs from d2) vs INTT duos
  int key_moduli_id = 0;
  int count = 0;
  int maxz = 1;
  std::vector<UnevenCt> host;
  UnevenCt *dev = nullptr;
  uint32_t *units = nullptr;
  // per-ciphertext workspace offsets (int64 elements)
  std::vector<int64_t> off_tmp, off_digits, off_ext, off_acc;
  int64_t ws_elems = 0;
  struct Launch {
    int stage, variant, offset, count;
  };
  std::vector<Launch> launches;
  Schedule sched[kStCount];
  bool tuned = false;
};

void uneven_plan_free(UnevenPlan *plan) {
  if (!plan)
    return;
  (void)hipFree(plan->dev);
  (void)hipFree(plan->units);
  delete plan;
}

namespace {

int lazy_variant(const fhe::CkksParams &P, int prime) {
  return P.q[prime] < fhe::kLazySmallBound ? fhe::kLazySmall
                                           : fhe::kLazyLarge;
}

// Build (or find) the plan of a batch shape.
UnevenPlan *uneven_plan(fhe_context *ctx, bool relin,
                        const std::vector<int> &mids, int key_moduli_id) {
  std::vector<int> key = {relin ? 1 : 0, key_moduli_id};
  key.insert(key.end(), mids.begin(), mids.end());
  auto it = ctx->uneven_plans.find(key);
  if (it != ctx->uneven_plans.end())
    return it->second;
  ctx->require_not_capturing("uneven batch plan");
  const auto &P = ctx->params;
  const int count = int(mids.size());
  const int tiles = int(P.n >> 12);
  const int chunks = int(P.n / fhe::kBlock);
  auto plan = std::make_unique<UnevenPlan>();
  plan->relin = relin;
  plan->key_moduli_id = key_moduli_id;
  plan->count = count;
  plan->host.resize(count);
  int max_size = 1;
  for (int c = 0; c < count; c++) {
    const int m = mids[c];
    UnevenCt &u = plan->host[c];
    u = UnevenCt{};
    u.moduli_id = m;
    u.limbs = P.limbs_at(m);
    u.ext = u.limbs + P.alpha;
    const auto parts = P.partitions(m, key_moduli_id);
    require(int(parts.size()) <= fhe::kUnevenMaxParts,
            "Uneven batch: too many key-switching partitions");
    u.parts = int(parts.size());
    u.key_partition0 = parts[0].key_index;
    for (int p = 0; p < u.parts; p++) {
      u.start[p] = parts[p].start;
      u.size[p] = parts[p].size;
      if (!P.modup_non_negative(m, parts[p]))
        u.signed_mask |= 1u << p;
      max_size = std::max(max_size, parts[p].size);
    }
    u.cst = ctx->modup_constants(m, key_moduli_id);
    u.cst_stride = P.ks_stride(m);
    u.ext_shoup_offset = P.ks_ext_shoup_offset(m);
  }
  while (plan->maxz < max_size)
    plan->maxz <<= 1;
  // workspace: [trio copy (relin) or rotated-before-permutation duo]
  // [digits][extended limbs][accumulators] per ciphertext
  int64_t off = 0;
  for (int c = 0; c < count; c++) {
    const UnevenCt &u = plan->host[c];
    plan->off_tmp.push_back(off);
    off += int64_t(relin ? 3 : 2) * u.limbs;
    plan->off_digits.push_back(off);
    off += int64_t(u.parts) * plan->maxz;
    plan->off_ext.push_back(off);
    off += int64_t(u.parts) * u.ext;
    plan->off_acc.push_back(off);
    off += 2 * int64_t(u.ext);
  }
  plan->ws_elems = off * P.n;

  // work lists: per stage, grouped by arithmetic variant, each group sorted
  // by (absolute prime, tile) so that the units sharing a prime's twiddles
  // or a key tile are neighbours
  std::vector<uint32_t> all;
  struct Keyed {
    int64_t key;
    uint32_t unit;
  };
  auto emit = [&](int stage, std::vector<Keyed> (&by_variant)[4]) {
    for (int v = 0; v < 4; v++) {
      auto &list = by_variant[v];
      if (list.empty())
        continue;
      std::stable_sort(list.begin(), list.end(),
                       [](const Keyed &a, const Keyed &b) {
                         return a.key < b.key;
                       });
      plan->launches.push_back({stage, v, int(all.size()), int(list.size())});
      for (auto &k : list)
        all.push_back(k.unit);
    }
  };
  auto tile_key = [&](int prime, int tile) {
    return int64_t(prime) * 256 + tile;
  };
  if (relin) {
    // the trio's d2 exact (it feeds ModUp), d0 / d1 lazy
    for (int stage : {int(kStTrioIntt1), int(kStTrioIntt2)}) {
      std::vector<Keyed> lists[4];
      for (int c = 0; c < count; c++) {
        const UnevenCt &u = plan->host[c];
        for (int poly = 0; poly < 3; poly++)
          for (int l = 0; l < u.limbs; l++)
            for (int t = 0; t < tiles; t++) {
              const int prime = u.moduli_id + l;
              const int v = poly == 2 ? fhe::kExact : lazy_variant(P, prime);
              lists[v].push_back({tile_key(prime, t),
                                  fhe::unit_pack(c, poly, l, t)});
            }
      }
      emit(stage, lists);
    }
  }
  {
    std::vector<Keyed> lists[4];
    for (int c = 0; c < count; c++)
      for (int p = 0; p < plan->host[c].parts; p++)
        for (int ch = 0; ch < chunks; ch++)
          lists[0].push_back({0, fhe::unit_pack(c, p, 0, ch)});
    emit(kStDigits, lists);
  }
  {
    std::vector<Keyed> lists[4];
    for (int c = 0; c < count; c++) {
      const UnevenCt &u = plan->host[c];
      for (int p = 0; p < u.parts; p++)
        for (int i = 0; i < u.ext; i++) {
          // a relinearized trio's d2 is at hand in NTT form: its limbs are
          // the partition's own extended limbs
          if (relin && i >= u.start[p] && i < u.start[p] + u.size[p])
            continue;
          const int prime = u.moduli_id + i;
          const int v = ((u.signed_mask >> p) & 1) ? fhe::kExact
                                                   : lazy_variant(P, prime);
          for (int t = 0; t < tiles; t++)
            lists[v].push_back({tile_key(prime, t),
                                fhe::unit_pack(c, p, i, t)});
        }
    }
    emit(kStExtend, lists);
  }
  {
    std::vector<Keyed> lists[4];
    for (int c = 0; c < count; c++) {
      const UnevenCt &u = plan->host[c];
      for (int i = 0; i < u.ext; i++) {
        const int prime = u.moduli_id + i;
        int v = lazy_variant(P, prime);
        // small primes of ciphertexts with several partitions: 96-bit sums
        if (v == fhe::kLazySmall && ctx->ks_wide_accumulate && u.parts > 1 &&
            u.parts <= fhe::kMaxWideParts)
          v = 3;
        for (int t = 0; t < tiles; t++)
          lists[v].push_back({tile_key(prime, t), fhe::unit_pack(c, 0, i, t)});
      }
    }
    emit(kStKs, lists);
  }
  {
    std::vector<Keyed> lists[4];
    for (int c = 0; c < count; c++) {
      const UnevenCt &u = plan->host[c];
      for (int h = 0; h < 2; h++)
        for (int i = 0; i < u.ext; i++) {
          const int prime = u.moduli_id + i;
          for (int t = 0; t < tiles; t++)
            lists[lazy_variant(P, prime)].push_back(
                {tile_key(prime, t), fhe::unit_pack(c, h, i, t)});
        }
    }
    emit(kStAccIntt2, lists);
  }
  {
    std::vector<Keyed> lists[4];
    for (int c = 0; c < count; c++)
      for (int h = 0; h < 2; h++)
        for (int ch = 0; ch < chunks; ch++)
          lists[0].push_back({0, fhe::unit_pack(c, h, 0, ch)});
    emit(kStModDown, lists);
  }
  if (!relin) {
    std::vector<Keyed> lists[4];
    const int aut_chunks = int(P.n / (4 * fhe::kBlock));
    for (int c = 0; c < count; c++)
      for (int poly = 0; poly < 2; poly++)
        for (int l = 0; l < plan->host[c].limbs; l++)
          for (int ch = 0; ch < aut_chunks; ch++)
            lists[0].push_back({0, fhe::unit_pack(c, poly, l, ch)});
    emit(kStAutomorphism, lists);
  }
  for (int s = 0; s < kStCount; s++)
    plan->sched[s] = kCandidates[0];

  hip_check(hipMalloc(&plan->dev, sizeof(UnevenCt) * count),
            "uneven plan allocation");
  hip_check(hipMemcpy(plan->dev, plan->host.data(), sizeof(UnevenCt) * count,
                      hipMemcpyHostToDevice),
            "uneven plan upload");
  hip_check(hipMalloc(&plan->units, sizeof(uint32_t) * all.size()),
            "uneven plan allocation");
  hip_check(hipMemcpy(plan->units, all.data(), sizeof(uint32_t) * all.size(),
                      hipMemcpyHostToDevice),
            "uneven plan upload");
  UnevenPlan *raw = plan.release();
  ctx->uneven_plans[key] = raw;
  return raw;
}

// One key switch of the plan's batch: relinearization of NTT-form trios
// (src[c] [3][limbs][N]) into INTT-form duos, or a rotation of INTT-form
// duos (src[c] [2][limbs][N]) by the automorphism g (g = 0: a plain key
// switch, CKKSCore::switch_key) into out[c].
void uneven_run(fhe_context *ctx, UnevenPlan *plan, const int64_t *const *src,
                int64_t *const *out, const int64_t *key, uint32_t g) {
  const auto &P = ctx->params;
  const int64_t N = P.n;
  const auto &t = ctx->tables;
  int64_t *ws = ctx->scratch(size_t(plan->ws_elems) * 8);
  UnevenPtrs in{}, tmp{}, digits{}, ext{}, acc{}, own{}, x{}, dst{}, c0{};
  for (int c = 0; c < plan->count; c++) {
    const int limbs = plan->host[c].limbs;
    in.p[c] = src[c];
    tmp.p[c] = ws + plan->off_tmp[c] * N;
    digits.p[c] = ws + plan->off_digits[c] * N;
    ext.p[c] = ws + plan->off_ext[c] * N;
    acc.p[c] = ws + plan->off_acc[c] * N;
    dst.p[c] = out[c];
    if (plan->relin) {
      own.p[c] = src[c] + 2 * limbs * N;      // NTT-form d2
      x.p[c] = tmp.p[c] + 2 * limbs * N;      // INTT-form d2
    } else {
      own.p[c] = nullptr;
      x.p[c] = src[c] + limbs * N;            // c1
      c0.p[c] = src[c];
    }
  }
  const int key_chain = P.limbs_at(plan->key_moduli_id) + P.alpha;
  for (const auto &L : plan->launches) {
    const Schedule s = plan->sched[L.stage];
    const UnitList list{plan->units + L.offset, L.count, s.xcd};
    ctx->stage(kStageNames[L.stage], [&] {
      switch (L.stage) {
      case kStTrioIntt1:
        fhe::launch_uneven_ntt(t, plan->dev, list, in, tmp, 0, true, 0,
                               nullptr, fhe::kInttRaw, false, L.variant,
                               s.cap, ctx->stream);
        break;
      case kStTrioIntt2:
        fhe::launch_uneven_ntt(t, plan->dev, list, tmp, tmp, 0, true, 8,
                               nullptr, fhe::kInttExit, true, L.variant,
                               s.cap, ctx->stream);
        break;
      case kStDigits:
        fhe::launch_uneven_digits(t, plan->dev, list, x, digits, plan->maxz,
                                  s.cap, ctx->stream);
        break;
      case kStExtend:
        fhe::launch_uneven_extend(t, plan->dev, list, digits, ext,
                                  plan->maxz, L.variant, s.cap, ctx->stream);
        break;
      case kStKs:
        fhe::launch_uneven_ks(t, plan->dev, list, ext, own, acc, key,
                              key_chain, plan->key_moduli_id,
                              ctx->ks_lazy_accumulate, L.variant, s.cap,
                              ctx->stream);
        break;
      case kStAccIntt2:
        fhe::launch_uneven_ntt(t, plan->dev, list, acc, acc, 1, true, 8,
                               nullptr, fhe::kInttExitReduce, true, L.variant,
                               s.cap, ctx->stream);
        break;
      case kStModDown:
        if (plan->relin)
          // y_h = d_h + ModDown(acc_h) with the INTT-form d0, d1
          fhe::launch_uneven_moddown(t, plan->dev, list, acc, tmp, 2, dst,
                                     s.cap, ctx->stream);
        else if (g == 0)
          fhe::launch_uneven_moddown(t, plan->dev, list, acc, c0, 1, dst,
                                     s.cap, ctx->stream);
        else
          // c0 joins after ModDown, where the reference adds c0 * P before
          // it (src/ckks/core_key_switch.cpp:509-548): both give
          // floor((X + c0 P) / P) = floor(X / P) + c0
          fhe::launch_uneven_moddown(t, plan->dev, list, acc, c0, 1, tmp,
                                     s.cap, ctx->stream);
        break;
      case kStAutomorphism:
        if (g != 0)
          fhe::launch_uneven_automorphism(t, plan->dev, list, tmp, dst, g,
                                          s.cap, ctx->stream);
        break;
      }
    });
  }
  ctx->check_launch("Uneven key switch");
}

// Choose each stage's schedule by timing the whole key switch (every stage
// reads what the stages before it wrote, and some work in place, so a stage
// is not rerun alone): one pass of coordinate descent from the default.  The
// trial runs write into a buffer of their own, so a caller's output that
// aliases its input (a rotation in place) is only written by the real call.
void uneven_tune(fhe_context *ctx, UnevenPlan *plan, const int64_t *const *src,
                 const int64_t *key, uint32_t g) {
  const int64_t N = ctx->params.n;
  int64_t total = 0;
  for (const auto &u : plan->host)
    total += 2 * int64_t(u.limbs) * N;
  void *trial = nullptr;
  hip_check(hipMalloc(&trial, size_t(total) * 8), "autotune allocation");
  std::vector<int64_t *> out(plan->count);
  int64_t off = 0;
  for (int c = 0; c < plan->count; c++) {
    out[c] = static_cast<int64_t *>(trial) + off;
    off += 2 * plan->host[c].limbs * N;
  }
  hipEvent_t a = nullptr, b = nullptr;
  hip_check(hipEventCreate(&a), "hipEventCreate");
  hip_check(hipEventCreate(&b), "hipEventCreate");
  const bool timing = ctx->timing;
  ctx->timing = false;
  auto measure = [&] {
    float best = 1e30f;
    for (int rep = 0; rep < 2; rep++) {
      hip_check(hipEventRecord(a, ctx->stream), "hipEventRecord");
      uneven_run(ctx, plan, src, out.data(), key, g);
      hip_check(hipEventRecord(b, ctx->stream), "hipEventRecord");
      hip_check(hipEventSynchronize(b), "hipEventSynchronize");
      float ms = 0;
      hip_check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
      best = std::min(best, ms);
    }
    return best;
  };
  uneven_run(ctx, plan, src, out.data(), key, g); // warm caches, workspace
  float best = measure();
  for (int s = 0; s < kStCount; s++) {
    bool used = false;
    for (const auto &L : plan->launches)
      used = used || L.stage == s;
    if (!used)
      continue;
    Schedule keep = plan->sched[s];
    for (int k = 0; k < kCandidateCount; k++) {
      if (kCandidates[k].cap == keep.cap && kCandidates[k].xcd == keep.xcd)
        continue;
      const Schedule prev = plan->sched[s];
      plan->sched[s] = kCandidates[k];
      const float ms = measure();
      if (ms < best * 0.99f) {
        best = ms;
        keep = kCandidates[k];
      }
      plan->sched[s] = prev;
    }
    plan->sched[s] = keep;
  }
  plan->tuned = true;
  ctx->timing = timing;
  hip_check(hipStreamSynchronize(ctx->stream), "autotune sync");
  (void)hipFree(trial);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
}

void check_uneven(fhe_context *ctx, int count, const int *moduli_ids,
                  int key_moduli_id, const char *op) {
  require(count > 0, std::string(op) + ": empty batch");
  require(fhe::ks_fused_supported(ctx->tables),
          std::string(op) + ": uneven-level batches need N = 2^16");
  for (int c = 0; c < count; c++) {
    ctx->check_moduli(moduli_ids[c], 1, op);
    ctx->check_key(moduli_ids[c], key_moduli_id, op);
  }
}

// The key switch of `count` ciphertexts in groups of at most kUnevenMax.
void uneven_switch(fhe_context *ctx, bool relin, int count,
                   const int *moduli_ids, const int64_t *const *src,
                   int64_t *const *out, const int64_t *key,
                   int key_moduli_id, uint32_t g) {
  for (int c0 = 0; c0 < count; c0 += fhe::kUnevenMax) {
    const int n = std::min(fhe::kUnevenMax, count - c0);
    std::vector<int> mids(moduli_ids + c0, moduli_ids + c0 + n);
    UnevenPlan *plan = uneven_plan(ctx, relin, mids, key_moduli_id);
    if (ctx->uneven_autotune && !plan->tuned && !ctx->capturing)
      uneven_tune(ctx, plan, src + c0, key, g);
    uneven_run(ctx, plan, src + c0, out + c0, key, g);
  }
}

} // namespace

struct fhe_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  fhe_context *ctx = nullptr;
};

extern "C" {

int fhe_relinearize_uneven(fhe_context *ctx, int count,
                           const int *moduli_ids, const int64_t *const *trios,
                           const int64_t *key, int key_moduli_id,
                           int64_t *const *outs) {
  return guarded([&] {
    require(ctx && moduli_ids && trios && key && outs, "null argument");
    check_uneven(ctx, count, moduli_ids, key_moduli_id, "Relinearize");
    for (int c = 0; c < count; c++) {
      require(trios[c] && outs[c], "null argument");
      require(static_cast<const void *>(trios[c]) !=
                  static_cast<const void *>(outs[c]),
              "Relinearize: an output must not alias its input");
    }
    uneven_switch(ctx, true, count, moduli_ids, trios, outs, key,
                  key_moduli_id, 0);
  });
}

int fhe_rotate_uneven(fhe_context *ctx, int count, const int *moduli_ids,
                      const int64_t *const *ins, const int64_t *key,
                      int key_moduli_id, int delta, int64_t *const *outs) {
  return guarded([&] {
    require(ctx && moduli_ids && ins && key && outs, "null argument");
    check_uneven(ctx, count, moduli_ids, key_moduli_id, "Rotate");
    for (int c = 0; c < count; c++)
      require(ins[c] && outs[c], "null argument");
    uneven_switch(ctx, false, count, moduli_ids, ins, outs, key,
                  key_moduli_id,
                  automorphism_exponent(ctx->params.log_n, delta));
  });
}

int fhe_rotate_rotation_key_uneven(fhe_context *ctx, int count,
                                   const int *moduli_ids,
                                   const int64_t *const *ins, int key_count,
                                   const int64_t *const *keys,
                                   const int *key_moduli_ids,
                                   const int *key_deltas, const int *deltas,
                                   int64_t *const *outs) {
  return guarded([&] {
    require(ctx && moduli_ids && ins && keys && key_moduli_ids &&
                key_deltas && deltas && outs,
            "null argument");
    require(key_count > 0, "Rotate: empty rotation key");
    const auto &P = ctx->params;
    const int64_t N = P.n;
    for (int k = 0; k < key_count; k++) {
      require(keys[k] != nullptr, "null argument");
      check_uneven(ctx, count, moduli_ids, key_moduli_ids[k], "Rotate");
    }
    for (int c = 0; c < count; c++)
      require(ins[c] && outs[c], "null argument");
    // CKKSEngine::to_normalized_delta (src/ckks/engine.cpp:227-235)
    const int slots = P.slot_count();
    std::vector<int> norm(count);
    for (int c = 0; c < count; c++)
      norm[c] = ((deltas[c] % slots) + slots) % slots;
    // each fixed key in the key's order rotates the ciphertexts whose delta
    // has its bit (src/ckks/engine_rotate.cpp:142-170); a ciphertext's first
    // rotation reads its input, later ones its output in place
    std::vector<bool> started(count, false);
    for (int k = 0; k < key_count; k++) {
      std::vector<int> mids;
      std::vector<const int64_t *> src;
      std::vector<int64_t *> dst;
      for (int c = 0; c < count; c++) {
        if ((norm[c] & key_deltas[k]) == 0)
          continue;
        mids.push_back(moduli_ids[c]);
        src.push_back(started[c] ? outs[c] : ins[c]);
        dst.push_back(outs[c]);
        started[c] = true;
      }
      if (mids.empty())
        continue;
      uneven_switch(ctx, false, int(mids.size()), mids.data(), src.data(),
                    dst.data(), keys[k], key_moduli_ids[k],
                    automorphism_exponent(P.log_n, key_deltas[k]));
    }
    // delta 0 returns the INTT-form input unchanged (engine_rotate.cpp:124)
    for (int c = 0; c < count; c++)
      if (!started[c] && outs[c] != ins[c])
        hip_check(hipMemcpyAsync(outs[c], ins[c],
                                 size_t(2 * P.limbs_at(moduli_ids[c])) * N * 8,
                                 hipMemcpyDeviceToDevice, ctx->stream),
                  "copy");
    ctx->check_launch("Rotate");
  });
}

int fhe_uneven_schedule(fhe_context *ctx, int relin, int count,
                        const int *moduli_ids, int key_moduli_id, int *caps,
                        int *xcds, int capacity) {
  return guarded([&] {
    require(ctx && moduli_ids && caps && xcds, "null argument");
    require(capacity >= kStCount, "Uneven schedule: capacity below the "
                                  "stage count");
    std::vector<int> key = {relin ? 1 : 0, key_moduli_id};
    key.insert(key.end(), moduli_ids, moduli_ids + count);
    auto it = ctx->uneven_plans.find(key);
    require(it != ctx->uneven_plans.end(),
            "Uneven schedule: no batch of this shape has run");
    for (int s = 0; s < kStCount; s++) {
      caps[s] = it->second->sched[s].cap;
      xcds[s] = it->second->sched[s].xcd;
    }
  });
}

int fhe_graph_begin(fhe_context *ctx) {
  return guarded([&] {
    require(ctx != nullptr, "null argument");
    require(!ctx->capturing, "Graph: the stream is already captured");
    require(!ctx->timing, "Graph: stage timing is on");
    hip_check(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed),
              "hipStreamBeginCapture");
    ctx->capturing = true;
  });
}

int fhe_graph_end(fhe_context *ctx, fhe_graph **out) {
  return guarded([&] {
    require(ctx && out, "null argument");
    require(ctx->capturing, "Graph: the stream is not captured");
    auto g = std::make_unique<fhe_graph>();
    ctx->capturing = false;
    hip_check(hipStreamEndCapture(ctx->stream, &g->graph),
              "hipStreamEndCapture");
    hip_check(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0),
              "hipGraphInstantiate");
    g->ctx = ctx;
    ctx->live_graphs++;
    *out = g.release();
  });
}

int fhe_graph_launch(fhe_context *ctx, fhe_graph *graph) {
  return guarded([&] {
    require(ctx && graph, "null argument");
    require(graph->ctx == ctx, "Graph: recorded on another context");
    require(!ctx->capturing, "Graph: the stream is being captured");
    hip_check(hipGraphLaunch(graph->exec, ctx->stream), "hipGraphLaunch");
  });
}

void fhe_graph_destroy(fhe_graph *graph) {
  if (!graph)
    return;
  if (graph->ctx && graph->ctx->live_graphs > 0)
    graph->ctx->live_graphs--;
  if (graph->exec)
    (void)hipGraphExecDestroy(graph->exec);
  if (graph->graph)
    (void)hipGraphDestroy(graph->graph);
  delete graph;
}

int fhe_context_uneven_autotune(fhe_context *ctx, int on) {
  return guarded([&] {
    require(ctx != nullptr, "null argument");
    ctx->uneven_autotune = on != 0;
  });
}

} // extern "C"
