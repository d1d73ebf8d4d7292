// fhe_context: device-resident CKKS parameters, tables, workspace and the
// key-switching pipelines shared by the C ABI's source files (context.hip:
// context, memory, transforms and the metric's pipeline; engine.hip: the
// engine-level compositions and the bootstrap; batch.hip: mixed-level
// batches; wire.hip: the wire format).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <complex>
#include <tuple>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "fhe_amd.h"
#include "kernels.h"
#include "params.h"

namespace fhe_host {

// the calling thread's last error message (fhe_last_error)
extern thread_local std::string g_last_error;

struct FheError : std::runtime_error {
  int code;
  FheError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char *what) {
  if (e == hipErrorOutOfMemory)
    throw FheError(FHE_ERR_OUT_OF_MEMORY, std::string(what) + ": out of memory");
  if (e != hipSuccess)
    throw FheError(FHE_ERR_DEVICE,
                   std::string(what) + ": " + hipGetErrorString(e));
}

inline void require(bool ok, const std::string &message) {
  if (!ok)
    throw FheError(FHE_ERR_INVALID, message);
}

// Exponent g of the automorphism X -> X^g for a rotation by delta slots:
// 5^-delta mod 2N, or -1 for conjugation (delta = -1 mod 2N); the
// reference's get_jump (src/ckks/encode.cpp:513-526).
inline uint32_t automorphism_exponent(int log_n, int delta) {
  const int64_t two_n = int64_t(2) << log_n;
  const int64_t shift = ((int64_t(delta) % two_n) + two_n) % two_n;
  if (shift == two_n - 1)
    return uint32_t(shift);
  int64_t g = 1, b = 5;
  for (int64_t e = two_n - shift; e > 0; e >>= 1) {
    if (e & 1)
      g = (g * b) & (two_n - 1);
    b = (b * b) & (two_n - 1);
  }
  return uint32_t(g);
}

template <class F> int guarded(F &&f) {
  try {
    f();
    return FHE_OK;
  } catch (const FheError &e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::invalid_argument &e) {
    g_last_error = e.what();
    return FHE_ERR_INVALID;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return FHE_ERR_DEVICE;
  }
}

template <class T> T *upload(const std::vector<T> &v, std::vector<void *> &own) {
  void *p = nullptr;
  hip_check(hipMalloc(&p, std::max<size_t>(1, v.size() * sizeof(T))),
            "table allocation");
  own.push_back(p);
  if (!v.empty())
    hip_check(hipMemcpy(p, v.data(), v.size() * sizeof(T),
                        hipMemcpyHostToDevice),
              "table upload");
  return static_cast<T *>(p);
}

// A C ABI status as an exception (for compositions of entry points).
inline void ok(int rc) {
  if (rc != FHE_OK)
    throw FheError(rc, fhe_last_error());
}

} // namespace fhe_host

using namespace fhe_host;

// Work lists of one uneven-level batch shape (uneven.hip).
struct UnevenPlan;
void uneven_plan_free(UnevenPlan *plan);

struct fhe_context {
  fhe::CkksParams params;
  fhe::DeviceTables tables{};
  hipStream_t stream = nullptr;
  int device = 0;
  std::vector<void *> owned;
  // per-(moduli_id, key_moduli_id) ModUp constants on the device
  std::map<std::pair<int, int>, int64_t *> modup_cache;
  // growable scratch for the fused pipelines
  void *workspace = nullptr;
  size_t workspace_bytes = 0;
  // lazy transforms where results are canonicalized before use
  bool lazy_key_switch = true;
  // ModUp fused with the first NTT pass of the extended limbs
  bool fuse_modup = false;
  // input NTT second pass + tensor + trio INTT first pass in one kernel
  bool fuse_tensor = false;
  // ModUp staged through the Garner digits (digits kernel + extension fused
  // with the first NTT pass)
  bool staged_modup = true;
  // each partition's own extended limbs from the NTT-form d2 when at hand
  bool own_limbs = true;
  // the coefficient forms of d0 and d1 on a second stream, concurrent with
  // the key switching of d2 (a second stream and two events, default
  // priority)
  // (1: forked once the trio is written; 2: once d2's INTT is done, so
  // that the side stream runs beside the key inner product)
  int overlap_trio = 1;
  hipStream_t side = nullptr;
  hipEvent_t fork_event = nullptr, join_event = nullptr;
  // small-prime key products accumulated without guards, reduced once
  bool ks_lazy_accumulate = true;
  // small-prime key products summed unreduced (96 bits), reduced once:
  // about 10% fewer VALU instructions in the key inner product, but it
  // issues them at a lower rate (measured slower), so off by default
  bool ks_wide_accumulate = false;
  // key inner product loads the next partition's tile ahead
  bool ks_prefetch = false;
  // waves per SIMD targeted by the key inner product's register allocation
  int ks_occupancy = 2;
  // second growable device buffer (Garner digits of the staged ModUp)
  void *aux = nullptr;
  size_t aux_bytes = 0;
  // third growable buffer: intermediates of composite operations whose steps
  // use the other two
  void *stash_buf = nullptr;
  size_t stash_bytes = 0;
  // encoding tables, uploaded on the first encode
  fhe::EncodeDevice encode{};
  bool encode_ready = false;
  // decoding tables, uploaded on the first decode; CRT constants per moduli_id
  fhe::DecodeDevice decode{};
  bool decode_ready = false;
  std::map<int, fhe::DecodeCrtDevice> decode_crt_cache;
  // device buffers of the engine-level compositions, kept by size for
  // reuse by later calls
  std::multimap<size_t, void *> text_pool;
  // prepared (rescaled, NTT-form) plaintexts of every diagonal of one
  // butterfly stage, resident per (direction, stages, stage, level)
  struct StagePlaintexts {
    void *data = nullptr;
    int count = 0, level = 0;
  };
  std::map<std::tuple<bool, int, int, int>, StagePlaintexts> boot_diagonals;
  // bootstrap slot-transform diagonals per (log slots, stages, direction)
  std::map<std::tuple<int, int, bool>,
           std::vector<std::vector<std::vector<std::complex<double>>>>>
      boot_twiddles;
  // uneven-level batches: work lists per (kind, key moduli_id, moduli_ids)
  std::map<std::vector<int>, UnevenPlan *> uneven_plans;
  // launch schedules of the uneven-level kernels chosen by timing the
  // candidates on the first call of each batch shape
  bool uneven_autotune = false;
  // the context's stream is being captured into a HIP graph: no call may
  // synchronize, allocate or upload
  bool capturing = false;

  void *pool_take(size_t bytes) {
    auto it = text_pool.find(bytes);
    if (it != text_pool.end()) {
      void *p = it->second;
      text_pool.erase(it);
      return p;
    }
    require_not_capturing("text allocation");
    void *p = nullptr;
    hip_check(hipMalloc(&p, bytes), "text allocation");
    return p;
  }
  // stage timing: HIP events recorded on `stream` around each stage while
  // enabled, folded into per-stage totals when read
  bool timing = false;
  struct PendingStage {
    const char *name;
    hipEvent_t start, stop;
  };
  std::vector<PendingStage> pending;
  std::vector<std::string> stage_names;
  std::vector<double> stage_ms;
  std::vector<int64_t> stage_launches;

  explicit fhe_context(fhe::CkksParams p) : params(std::move(p)) {
    hip_check(hipGetDevice(&device), "hipGetDevice");
    hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking),
              "stream creation");
    hip_check(hipStreamCreateWithFlags(&side, hipStreamNonBlocking),
              "stream creation");
    hip_check(hipEventCreateWithFlags(&fork_event, hipEventDisableTiming),
              "event creation");
    hip_check(hipEventCreateWithFlags(&join_event, hipEventDisableTiming),
              "event creation");
    auto &t = tables;
    t.q = upload(params.q, owned);
    t.k = upload(params.k, owned);
    t.q2 = upload(params.q2, owned);
    t.qh = upload(params.qh, owned);
    t.r2 = upload(params.r2, owned);
    t.ninv = upload(params.ninv, owned);
    t.root = upload(params.root, owned);
    t.root_inverse = upload(params.root_inverse, owned);
    t.root_shoup = upload(params.root_shoup, owned);
    t.root_inverse_shoup = upload(params.root_inverse_shoup, owned);
    t.p_inverse = upload(params.p_inverse_mult_r, owned);
    t.p_inverse_shoup = upload(params.p_inverse_shoup, owned);
    t.q_top = upload(params.q_top, owned);
    t.rescale = upload(params.rescale_scale, owned);
    t.host_q = params.q.data();
    t.split_variants = 1;
    t.prime_count = params.prime_count;
    t.log_n = params.log_n;
    t.alpha = params.alpha;
  }

  // teardown: errors are not actionable here, so they are ignored
  ~fhe_context() {
    if (stream)
      (void)hipStreamSynchronize(stream);
    if (side)
      (void)hipStreamSynchronize(side);
    for (auto &st : pending) {
      (void)hipEventDestroy(st.start);
      (void)hipEventDestroy(st.stop);
    }
    for (auto &kv : modup_cache)
      (void)hipFree(kv.second);
    for (void *p : owned)
      (void)hipFree(p);
    for (auto &kv : text_pool)
      (void)hipFree(kv.second);
    for (auto &kv : boot_diagonals)
      (void)hipFree(kv.second.data);
    for (auto &kv : uneven_plans)
      uneven_plan_free(kv.second);
    if (workspace)
      (void)hipFree(workspace);
    if (aux)
      (void)hipFree(aux);
    if (stash_buf)
      (void)hipFree(stash_buf);
    if (side)
      (void)hipStreamDestroy(side);
    if (fork_event)
      (void)hipEventDestroy(fork_event);
    if (join_event)
      (void)hipEventDestroy(join_event);
    if (stream)
      (void)hipStreamDestroy(stream);
  }

  int64_t n() const { return params.n; }

  // Runs f (kernel launches on `stream`), bracketed by events when timing.
  template <class F> void stage(const char *name, F &&f) {
    if (!timing) {
      f();
      return;
    }
    PendingStage st{name, nullptr, nullptr};
    hip_check(hipEventCreate(&st.start), "hipEventCreate");
    hip_check(hipEventCreate(&st.stop), "hipEventCreate");
    pending.push_back(st);
    hip_check(hipEventRecord(st.start, stream), "hipEventRecord");
    f();
    hip_check(hipEventRecord(st.stop, stream), "hipEventRecord");
  }

  void fold_timings() {
    hip_check(hipStreamSynchronize(stream), "timing sync");
    for (auto &st : pending) {
      float ms = 0;
      hip_check(hipEventElapsedTime(&ms, st.start, st.stop),
                "hipEventElapsedTime");
      size_t i = 0;
      while (i < stage_names.size() && stage_names[i] != st.name)
        i++;
      if (i == stage_names.size()) {
        stage_names.push_back(st.name);
        stage_ms.push_back(0);
        stage_launches.push_back(0);
      }
      stage_ms[i] += ms;
      stage_launches[i]++;
      (void)hipEventDestroy(st.start);
      (void)hipEventDestroy(st.stop);
    }
    pending.clear();
  }

  // extended polynomials held at once by key switching: every partition on
  // the fused path, one at a time otherwise
  int ext_slots(int moduli_id) const {
    return fhe::ks_fused_supported(tables) ? params.partition_count(moduli_id)
                                           : 1;
  }

  void require_not_capturing(const char *what) const {
    require(!capturing, std::string(what) +
                            ": not allowed while the stream is captured (run "
                            "the call once before capturing it)");
  }
  // captured graphs alive on this context: they hold the addresses of the
  // workspaces, which may then not move
  int live_graphs = 0;
  void require_workspace_fixed(const char *what) const {
    require_not_capturing(what);
    require(live_graphs == 0,
            std::string(what) + ": a captured graph holds the workspace "
                                "(destroy it, or run the larger call before "
                                "capturing)");
  }

  int64_t *aux_scratch(size_t bytes) {
    if (bytes > aux_bytes) {
      require_workspace_fixed("workspace growth");
      hip_check(hipStreamSynchronize(stream), "workspace sync");
      if (aux)
        hip_check(hipFree(aux), "workspace free");
      aux = nullptr;
      aux_bytes = 0;
      hip_check(hipMalloc(&aux, bytes), "workspace allocation");
      aux_bytes = bytes;
    }
    return static_cast<int64_t *>(aux);
  }

  int64_t *stash(size_t bytes) {
    if (bytes > stash_bytes) {
      require_workspace_fixed("workspace growth");
      hip_check(hipStreamSynchronize(stream), "workspace sync");
      if (stash_buf)
        hip_check(hipFree(stash_buf), "workspace free");
      stash_buf = nullptr;
      stash_bytes = 0;
      hip_check(hipMalloc(&stash_buf, bytes), "workspace allocation");
      stash_bytes = bytes;
    }
    return static_cast<int64_t *>(stash_buf);
  }

  // CKKSCore::switch_key (src/ckks/core.cpp:1058-1083) of `count` INTT-form
  // duos at moduli_id: key switching of c1, then c0 added to the first half
  // and the second half copied.  Uses the scratch workspace.
  void switch_key_into(int count, int moduli_id, const int64_t *in,
                       const int64_t *key, int key_moduli_id, int64_t *out,
                       bool signed_input = false) {
    const int limbs = params.limbs_at(moduli_id);
    const int ext_limbs = limbs + params.alpha;
    const int64_t N = n();
    const int64_t ext_l = int64_t(count) * ext_slots(moduli_id) * ext_limbs;
    const int64_t acc_l = int64_t(count) * 2 * ext_limbs;
    int64_t *ws = scratch(size_t(ext_l + acc_l) * N * 8);
    int64_t *ext = ws, *acc = ext + ext_l * N;
    key_switch(count, moduli_id, key, key_moduli_id, in + limbs * N,
               2 * limbs * N, ext, acc, nullptr, 0, signed_input);
    stage("moddown_add", [&] {
      fhe::launch_moddown_add(tables, acc, in, 2 * limbs * N, 1, out, count,
                              limbs, moduli_id, stream);
    });
  }

  int64_t *scratch(size_t bytes) {
    if (bytes > workspace_bytes) {
      require_workspace_fixed("workspace growth");
      hip_check(hipStreamSynchronize(stream), "workspace sync");
      if (workspace)
        hip_check(hipFree(workspace), "workspace free");
      workspace = nullptr;
      workspace_bytes = 0;
      hip_check(hipMalloc(&workspace, bytes), "workspace allocation");
      workspace_bytes = bytes;
    }
    return static_cast<int64_t *>(workspace);
  }

  const fhe::EncodeDevice &encode_tables() {
    if (!encode_ready) {
      std::vector<double> fc, fs, tc, ts;
      params.encode_trig(fc, fs, tc, ts);
      encode.pre = upload(params.encode_pre_permutation(), owned);
      encode.fft_cos = upload(fc, owned);
      encode.fft_sin = upload(fs, owned);
      encode.twist_cos = upload(tc, owned);
      encode.twist_sin = upload(ts, owned);
      encode.log_codec = params.log_n;
      encode_ready = true;
    }
    return encode;
  }

  const fhe::DecodeDevice &decode_tables() {
    if (!decode_ready) {
      std::vector<double> sc, ss, ic, is;
      params.decode_trig(sc, ss, ic, is);
      decode.post = upload(params.decode_post_permutation(), owned);
      decode.skew_cos = upload(sc, owned);
      decode.skew_sin = upload(ss, owned);
      decode.ifft_cos = upload(ic, owned);
      decode.ifft_sin = upload(is, owned);
      decode.log_codec = params.log_n;
      decode_ready = true;
    }
    return decode;
  }

  const fhe::DecodeCrtDevice &decode_crt(int moduli_id) {
    auto it = decode_crt_cache.find(moduli_id);
    if (it != decode_crt_cache.end())
      return it->second;
    const auto h = params.decode_crt(moduli_id);
    require(fhe::decode_crt_supported(h.words),
            "Decode: the modulus product is too wide");
    fhe::DecodeCrtDevice d{};
    d.reverser = upload(h.reverser, owned);
    d.q_product = upload(h.q_product, owned);
    d.q_half = upload(h.q_half, owned);
    d.upper = upload(h.upper, owned);
    d.k = h.k;
    d.words = h.words;
    return decode_crt_cache[moduli_id] = d;
  }

  // Level down of `polys` INTT-form polynomials at `level` (rescaled or not)
  // to dst_level < level; out is not rescaled at dst_level.
  void level_down_into(const int64_t *in, int64_t *out, int polys, int level,
                       bool rescaled, int dst_level) {
    const auto &P = params;
    const int dst_mid = P.moduli_id_of_level(dst_level);
    if (rescaled) {
      // CKKSCore::level_down_after_rescale (src/ckks/core.cpp:743-779): the
      // rescaled input sits at moduli_id_of_level(level) + 1
      const int src_mid = P.moduli_id_of_level(level) + 1;
      const auto c = P.level_down_multiplier_after_rescale(level, dst_level);
      const int limbs = int(c.size());
      require(limbs <= fhe::kMaxEwLimbs, "Level down: too many limbs");
      const int64_t src_poly = int64_t(P.limbs_at(src_mid)) << P.log_n;
      const int64_t dst_poly = int64_t(limbs) << P.log_n;
      const int64_t offset = int64_t(dst_mid - src_mid) << P.log_n;
      stage("elementwise", [&] {
        for (int p = 0; p < polys; p++)
          fhe::launch_elementwise(tables, fhe::kEwEnterReduce,
                                  in + p * src_poly + offset, nullptr,
                                  out + p * dst_poly, limbs, limbs, dst_mid,
                                  c.data(), stream);
      });
    } else {
      // CKKSCore::level_down (src/ckks/core.cpp:792-838)
      const int src_mid = P.moduli_id_of_level(level);
      const int next_mid = P.moduli_id_of_level(level - 1);
      const auto c = P.level_down_multiplier(level, dst_level);
      require(int(c.size()) <= fhe::kMaxEwLimbs, "Level down: too many limbs");
      stage("scalar_rescale", [&] {
        fhe::launch_scalar_rescale(tables, in, out, polys, int(c.size()),
                                   src_mid, c.data(), dst_mid - next_mid,
                                   stream);
      });
    }
  }

  // Chain position of a text at `level`: a rescaled text sits one modulus
  // above the not-rescaled text of its level (the not_rescaled_levels_ /
  // rescaled_levels_ maps, src/ckks/parameters.cpp:60-90).
  int moduli_id_of(int level, bool rescaled) const {
    return params.moduli_id_of_level(level) + (rescaled ? 1 : 0);
  }

  // The engine's form changes of `polys` polynomials at (level, flags), in
  // its order: intt_ first when asked (CKKSEngine::intt_), level_down_ to
  // dst_level <= level (src/ckks/engine_level_down.cpp:5-37: after rescale
  // in the current form, else in the coefficient domain), rescale_ when
  // `rescale` and not yet rescaled (src/ckks/engine_rescale.cpp:5-33), ntt_
  // when `ntt` (src/ckks/engine_ntt.cpp:5-21).  Returns `in` itself when
  // nothing changes, else `out`; t0 and t1 hold polys x limbs(level, flags)
  // x N each, and out may not alias them.
  const int64_t *transform_into(const int64_t *in, int64_t polys, int level,
                                int flags, bool intt_first, int dst_level,
                                bool rescale, bool ntt, int64_t *out,
                                int64_t *t0, int64_t *t1) {
    bool rescaled = (flags & FHE_FLAG_RESCALED) != 0;
    bool is_ntt = (flags & FHE_FLAG_NTT_FORM) != 0;
    const int64_t *cur = in;
    // the steps that write a result, counted first so that the last one
    // writes to out (the others alternate between t0 and t1)
    int steps = 0;
    {
      bool r = rescaled, nt = is_ntt;
      if (intt_first && nt) {
        steps++;
        nt = false;
      }
      if (level > dst_level) {
        steps += (!r && nt) ? 3 : 1;
        r = false;
      }
      if (rescale && !r)
        steps += nt ? 3 : 1;
      if (ntt && !nt)
        steps++;
    }
    int64_t *bufs[2] = {t0, t1};
    int next = 0, done = 0;
    auto target = [&] {
      if (++done == steps)
        return out;
      int64_t *b = bufs[next];
      next ^= 1;
      return b;
    };
    auto to_intt = [&] {
      const int mid = moduli_id_of(level, rescaled);
      const int limbs = params.limbs_at(mid);
      int64_t *y = target();
      stage("intt", [&] {
        fhe::ntt_inverse(tables, cur, y, polys * limbs, limbs, mid,
                         fhe::kInttExit, stream);
      });
      cur = y;
      is_ntt = false;
    };
    auto to_ntt = [&] {
      const int mid = moduli_id_of(level, rescaled);
      const int limbs = params.limbs_at(mid);
      int64_t *y = target();
      stage("ntt", [&] {
        fhe::ntt_forward(tables, cur, y, polys * limbs, limbs, mid, tables.r2,
                         stream);
      });
      cur = y;
      is_ntt = true;
    };
    if (intt_first && is_ntt)
      to_intt();
    if (level > dst_level) {
      // a not-rescaled NTT-form text goes down in the coefficient domain
      // and comes back
      const bool back = !rescaled && is_ntt;
      if (back)
        to_intt();
      int64_t *y = target();
      level_down_into(cur, y, int(polys), level, rescaled, dst_level);
      cur = y;
      rescaled = false;
      level = dst_level;
      if (back)
        to_ntt();
    }
    if (rescale && !rescaled) {
      require(level >= 1, "Rescale: the text should have a positive level.");
      const bool back = is_ntt;
      if (back)
        to_intt();
      const int mid = moduli_id_of(level, false);
      int64_t *y = target();
      stage("rescale", [&] {
        fhe::launch_rescale(tables, cur, y, int(polys), params.limbs_at(mid),
                            mid, stream);
      });
      cur = y;
      rescaled = true;
      if (back)
        to_ntt();
    }
    if (ntt && !is_ntt)
      to_ntt();
    require(done == steps && (steps == 0 || cur == out),
            "internal: form change steps miscounted");
    return steps == 0 ? in : out;
  }

  // CKKSEngine::prepare_for_multiply (src/ckks/engine_multiply.cpp:13-27):
  // intt_ when the level changes, level_down_, rescale_, ntt_.
  const int64_t *prepare_for_multiply_into(const int64_t *in, int64_t polys,
                                           int level, int flags,
                                           int dst_level, int64_t *out,
                                           int64_t *t0, int64_t *t1) {
    return transform_into(in, polys, level, flags, level > dst_level,
                          dst_level, true, true, out, t0, t1);
  }

  // Relinearization of `count` NTT-form trios at moduli_id into INTT-form
  // duos (CKKSCore::relinearize, src/ckks/core.cpp:713-741).  Uses the
  // scratch workspace; the trios are not modified.
  void relinearize_into(int count, int moduli_id, const int64_t *trios,
                        const int64_t *key, int key_moduli_id, int64_t *out) {
    const int limbs = params.limbs_at(moduli_id);
    const int ext_limbs = limbs + params.alpha;
    const int64_t N = n();
    const int64_t trio_l = int64_t(count) * 3 * limbs;
    const int64_t ext_l = int64_t(count) * ext_slots(moduli_id) * ext_limbs;
    const int64_t acc_l = int64_t(count) * 2 * ext_limbs;
    int64_t *ws = scratch(size_t(trio_l + ext_l + acc_l) * N * 8);
    int64_t *trio = ws, *ext = trio + trio_l * N, *acc = ext + ext_l * N;
    const fhe::LazySpec d01 =
        lazy_key_switch ? fhe::LazySpec{3, 2} : fhe::LazySpec();
    stage("intt(trio)", [&] {
      fhe::ntt_inverse(tables, trios, trio, trio_l, limbs, moduli_id,
                       fhe::kInttExit, stream, d01);
    });
    // the caller's NTT-form d2 provides the partitions' own extended limbs
    key_switch(count, moduli_id, key, key_moduli_id, trio + 2 * limbs * N,
               3 * limbs * N, ext, acc, trios + 2 * limbs * N, 3 * limbs * N);
    stage("moddown_relin", [&] {
      fhe::launch_moddown_relin(tables, acc, trio, out, count, limbs,
                                moduli_id, stream);
    });
  }

  const int64_t *modup_constants(int moduli_id, int key_moduli_id) {
    auto key = std::make_pair(moduli_id, key_moduli_id);
    auto it = modup_cache.find(key);
    if (it != modup_cache.end())
      return it->second;
    require_not_capturing("constant upload");
    const auto host = params.modup_constants(moduli_id, key_moduli_id);
    void *p = nullptr;
    hip_check(hipMalloc(&p, host.size() * sizeof(int64_t)), "constants");
    hip_check(hipMemcpy(p, host.data(), host.size() * sizeof(int64_t),
                        hipMemcpyHostToDevice),
              "constants upload");
    modup_cache[key] = static_cast<int64_t *>(p);
    return static_cast<int64_t *>(p);
  }

  void check_launch(const char *what) {
    hip_check(hipGetLastError(), what);
  }

  void check_moduli(int moduli_id, int min_limbs, const char *op) {
    require(moduli_id >= 0 && moduli_id < params.chain_count,
            std::string(op) + ": moduli_id out of range.");
    require(params.limbs_at(moduli_id) >= min_limbs,
            std::string(op) + ": the given level is too low.");
  }

  void check_key(int moduli_id, int key_moduli_id, const char *op) {
    require(key_moduli_id >= 0 && key_moduli_id <= moduli_id,
            std::string(op) +
                ": the key level is lower than the ciphertext level.");
    require(params.max_key_switching_level() >=
                params.level_of(key_moduli_id, false),
            std::string(op) + ": the key moduli id is above the key-switching "
                              "level.");
  }

  // Key switching of the INTT-form polynomial src + c * src_stride of each
  // ciphertext c into acc ([count][2][ext][N], INTT form, reduced).
  // own (optional): the NTT form of the same polynomials (limb i of
  // ciphertext c at own + c * own_stride + i * N); the fused path then takes
  // each partition's own extended limbs from it instead of transforming them.
  // signed_input: the source limbs may be any int64 (the signed tile of
  // modulus raising), not residues in [0, 2q): every partition then takes
  // the signed Garner digits and the exact extension.
  // digits (optional): room for the staged ModUp's Garner digits
  // (ks_digits_words) instead of the context's second growable buffer.
  void key_switch(int count, int moduli_id, const int64_t *key,
                  int key_moduli_id, const int64_t *src, int64_t src_stride,
                  int64_t *ext, int64_t *acc, const int64_t *own = nullptr,
                  int64_t own_stride = 0, bool signed_input = false,
                  int64_t *digits = nullptr) {
    if (fhe::ks_fused_supported(tables)) {
      if (!own_limbs || int(params.partitions(moduli_id, key_moduli_id)
                                .size()) > fhe::kMaxFusedParts)
        own = nullptr;
      ks_decompose(count, moduli_id, key_moduli_id, src, src_stride, ext,
                   own != nullptr, signed_input, digits);
      ks_apply(count, moduli_id, key, key_moduli_id, ext, acc, own,
               own_stride);
      return;
    }
    const int limbs = params.limbs_at(moduli_id);
    const int ext_limbs = limbs + params.alpha;
    const int64_t N = n();
    const int key_chain = params.limbs_at(key_moduli_id) + params.alpha;
    const int key_offset = params.limbs_at(key_moduli_id) - limbs;
    const auto parts = params.partitions(moduli_id, key_moduli_id);
    const int64_t *cst = modup_constants(moduli_id, key_moduli_id);
    const int stride = params.ks_stride(moduli_id);
    for (size_t p = 0; p < parts.size(); p++) {
      const auto &part = parts[p];
      fhe::launch_modup(tables, src, src_stride, ext, ext_limbs * N, count,
                        limbs, moduli_id, part.start, part.size,
                        cst + p * stride, params.ks_ext_shoup_offset(moduli_id),
                        !signed_input &&
                            params.modup_non_negative(moduli_id, part),
                        false, stream);
      fhe::ntt_forward(tables, ext, ext, int64_t(count) * ext_limbs,
                       ext_limbs, moduli_id, nullptr, stream);
      const int64_t *key0 =
          key + (int64_t(2 * part.key_index) * key_chain + key_offset) * N;
      const int64_t *key1 =
          key + (int64_t(2 * part.key_index + 1) * key_chain + key_offset) * N;
      fhe::launch_key_inner_product(tables, ext, ext_limbs * N, key0, key1,
                                    acc, 2 * ext_limbs * N, count, ext_limbs,
                                    moduli_id, p == 0, stream);
    }
    fhe::ntt_inverse(tables, acc, acc, int64_t(count) * 2 * ext_limbs,
                     ext_limbs, moduli_id, fhe::kInttExitReduce, stream);
  }

  // The key-independent half of key switching on the fused path (N = 2^16):
  // ModUp of every partition and the first forward NTT pass of every
  // extended limb into ext [count][parts][ext][N] (CKKSCore::decompose,
  // src/ckks/core_key_switch.cpp:434-460, up to its second NTT pass).
  // The extended polynomials, their products with the key and the
  // accumulators are only read by ModDown after the canonicalizing exit,
  // so with lazy_key_switch their arithmetic is lazy.
  // Words of Garner digits the staged ModUp of `count` ciphertexts at
  // moduli_id holds (partitions x partition size rounded to a power of 2).
  int64_t ks_digits_words(int count, int moduli_id) const {
    int maxz = 1;
    while (maxz < params.alpha)
      maxz <<= 1;
    return int64_t(count) * params.partition_count(moduli_id) * maxz *
           params.n;
  }

  void ks_decompose(int count, int moduli_id, int key_moduli_id,
                    const int64_t *src, int64_t src_stride, int64_t *ext,
                    bool skip_own = false, bool signed_input = false,
                    int64_t *digits_ws = nullptr) {
    const int limbs = params.limbs_at(moduli_id);
    const int ext_limbs = limbs + params.alpha;
    const int64_t N = n();
    const auto parts = params.partitions(moduli_id, key_moduli_id);
    const int64_t *cst = modup_constants(moduli_id, key_moduli_id);
    const int stride = params.ks_stride(moduli_id);
    const int np = int(parts.size());
    const fhe::LazySpec lazy = lazy_key_switch ? fhe::kAllLazy
                                               : fhe::LazySpec();
    int max_size = 0;
    bool is_signed[fhe::kMaxFusedParts] = {};
    std::vector<int> starts(np), sizes(np);
    for (int p = 0; p < np; p++) {
      starts[p] = parts[p].start;
      sizes[p] = parts[p].size;
      if (p < fhe::kMaxFusedParts)
        is_signed[p] = signed_input ||
                       !params.modup_non_negative(moduli_id, parts[p]);
      max_size = std::max(max_size, sizes[p]);
    }
    if (staged_modup && fhe::staged_modup_supported(tables, max_size, np)) {
      int maxz = 1;
      while (maxz < max_size)
        maxz <<= 1;
      int64_t *digits =
          digits_ws ? digits_ws
                    : aux_scratch(size_t(count) * np * maxz * size_t(N) * 8);
      stage("modup_staged(ext)", [&] {
        fhe::launch_staged_modup(tables, src, src_stride, digits, ext, count,
                                 limbs, moduli_id, starts.data(), sizes.data(),
                                 is_signed, np, cst, stride,
                                 params.ks_ext_shoup_offset(moduli_id),
                                 lazy_key_switch, skip_own, stream);
      });
      return;
    }
    if (fuse_modup && fhe::modup_ntt_fused_supported(tables, max_size, np)) {
      stage("modup_ntt_high(ext)", [&] {
        fhe::launch_modup_ntt_fused(tables, src, src_stride, ext, count, limbs,
                                    moduli_id, starts.data(), sizes.data(),
                                    is_signed, np, cst, stride,
                                    lazy_key_switch, stream);
      });
      return;
    }
    for (int p = 0; p < np; p++)
      stage("modup", [&] {
        fhe::launch_modup(tables, src, src_stride,
                          ext + int64_t(p) * ext_limbs * N,
                          int64_t(np) * ext_limbs * N, count, limbs, moduli_id,
                          parts[p].start, parts[p].size, cst + p * stride,
                          params.ks_ext_shoup_offset(moduli_id),
                          !signed_input &&
                              params.modup_non_negative(moduli_id, parts[p]),
                          lazy_key_switch && !signed_input, stream);
      });
    stage("ntt_pass_high(ext)", [&] {
      fhe::ntt_forward_first(tables, ext, ext, int64_t(count) * np * ext_limbs,
                             ext_limbs, moduli_id, nullptr, stream, lazy);
    });
  }

  // The key-dependent half on the fused path: second NTT pass, key inner
  // product over all partitions, both INTT passes (with exit and reduce)
  // into acc [count][2][ext][N].
  void ks_apply(int count, int moduli_id, const int64_t *key,
                int key_moduli_id, const int64_t *ext, int64_t *acc,
                const int64_t *own = nullptr, int64_t own_stride = 0) {
    const int limbs = params.limbs_at(moduli_id);
    const int ext_limbs = limbs + params.alpha;
    const int key_chain = params.limbs_at(key_moduli_id) + params.alpha;
    const int key_offset = params.limbs_at(key_moduli_id) - limbs;
    const auto parts = params.partitions(moduli_id, key_moduli_id);
    const int np = int(parts.size());
    const fhe::LazySpec lazy = lazy_key_switch ? fhe::kAllLazy
                                               : fhe::LazySpec();
    std::vector<int> starts(np), sizes(np);
    for (int p = 0; p < np; p++) {
      starts[p] = parts[p].start;
      sizes[p] = parts[p].size;
    }
    if (np > fhe::kMaxFusedParts)
      own = nullptr;
    stage("ks_inner_product_r16", [&] {
      fhe::launch_ks_inner_product_fused(
          tables, ext, key, key_chain, key_offset, parts[0].key_index, np, acc,
          count, ext_limbs, moduli_id, lazy_key_switch, own, own_stride,
          starts.data(), sizes.data(), ks_lazy_accumulate, ks_wide_accumulate,
          ks_prefetch, ks_occupancy, stream);
    });
    stage("intt_pass_high(acc)", [&] {
      fhe::ntt_inverse_second(tables, acc, acc, int64_t(count) * 2 * ext_limbs,
                              ext_limbs, moduli_id, fhe::kInttExitReduce,
                              stream, lazy);
    });
  }
};

// Device encoding (context.hip): `count` device messages at `level` into
// out [count][limbs (+ alpha with include_special)][N].
void encode_texts(fhe_context *ctx, const double *values, bool complex_input,
                  int count, int level, bool include_special, int64_t *out);

// The parameters a fhe_params describes (context.hip); twiddles = false
// skips the NTT tables (host-only queries).
fhe::CkksParams resolve_params(const fhe_params *p, bool twiddles = true);

// The batched HMult + relinearize + rescale pipeline (context.hip) on a
// caller-given workspace of hmult_workspace_words words and, when given, a
// digits buffer of ks_digits_words words.
int64_t hmult_workspace_words(const fhe_context *ctx, int count,
                              int moduli_id);
void hmult_relin_rescale_into(fhe_context *ctx, int count, int moduli_id,
                              const int64_t *a, const int64_t *b,
                              const int64_t *key, int key_moduli_id,
                              int64_t *out, int64_t *ws, int64_t *digits);
// Set-up of the allocation-free entry point on the current device `device`
// (noalloc.hip).
void noalloc_scratch_setup(int device);
