// fhe_context: device-resident CKKS parameters, workspace and the C ABI
// declared in include/fhe_amd.h.
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

namespace {

thread_local std::string g_last_error;

struct FheError : std::runtime_error {
  int code;
  FheError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char *what) {
  if (e == hipErrorOutOfMemory)
    throw FheError(FHE_ERR_OUT_OF_MEMORY, std::string(what) + ": out of memory");
  if (e != hipSuccess)
    throw FheError(FHE_ERR_DEVICE,
                   std::string(what) + ": " + hipGetErrorString(e));
}

void require(bool ok, const std::string &message) {
  if (!ok)
    throw FheError(FHE_ERR_INVALID, message);
}

// Exponent g of the automorphism X -> X^g for a rotation by delta slots:
// 5^-delta mod 2N, or -1 for conjugation (delta = -1 mod 2N); the
// reference's get_jump (src/ckks/encode.cpp:513-526).
uint32_t automorphism_exponent(int log_n, int delta) {
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

} // namespace

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
  bool overlap_trio = true;
  hipStream_t side = nullptr;
  hipEvent_t fork_event = nullptr, join_event = nullptr;
  // small-prime key products accumulated without guards, reduced once
  bool ks_lazy_accumulate = true;
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

  void *pool_take(size_t bytes) {
    auto it = text_pool.find(bytes);
    if (it != text_pool.end()) {
      void *p = it->second;
      text_pool.erase(it);
      return p;
    }
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

  int64_t *aux_scratch(size_t bytes) {
    if (bytes > aux_bytes) {
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
    const int64_t N = n();
    bool rescaled = (flags & FHE_FLAG_RESCALED) != 0;
    bool is_ntt = (flags & FHE_FLAG_NTT_FORM) != 0;
    const int64_t *cur = in;
    int64_t *bufs[2] = {t0, t1};
    int next = 0;
    auto target = [&] {
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
    if (cur == in)
      return in;
    const int limbs = params.limbs_at(moduli_id_of(level, rescaled));
    hip_check(hipMemcpyAsync(out, cur, size_t(polys * limbs) * N * 8,
                             hipMemcpyDeviceToDevice, stream),
              "copy");
    return out;
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
  void key_switch(int count, int moduli_id, const int64_t *key,
                  int key_moduli_id, const int64_t *src, int64_t src_stride,
                  int64_t *ext, int64_t *acc, const int64_t *own = nullptr,
                  int64_t own_stride = 0, bool signed_input = false) {
    if (fhe::ks_fused_supported(tables)) {
      if (!own_limbs || int(params.partitions(moduli_id, key_moduli_id)
                                .size()) > fhe::kMaxFusedParts)
        own = nullptr;
      ks_decompose(count, moduli_id, key_moduli_id, src, src_stride, ext,
                   own != nullptr, signed_input);
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
  void ks_decompose(int count, int moduli_id, int key_moduli_id,
                    const int64_t *src, int64_t src_stride, int64_t *ext,
                    bool skip_own = false, bool signed_input = false) {
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
          aux_scratch(size_t(count) * np * maxz * size_t(N) * 8);
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
          starts.data(), sizes.data(), ks_lazy_accumulate, ks_prefetch,
          ks_occupancy, stream);
    });
    stage("intt_pass_high(acc)", [&] {
      fhe::ntt_inverse_second(tables, acc, acc, int64_t(count) * 2 * ext_limbs,
                              ext_limbs, moduli_id, fhe::kInttExitReduce,
                              stream, lazy);
    });
  }
};

extern "C" {

const char *fhe_last_error(void) { return g_last_error.c_str(); }

static fhe::CkksParams resolve_params(const fhe_params *p,
                                      bool twiddles = true) {
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
    *out = new fhe_context(std::move(params));
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
      ctx->overlap_trio = value != 0;
    else if (n == "ks_occupancy")
      ctx->ks_occupancy = value;
    else if (n == "ks_prefetch")
      ctx->ks_prefetch = value != 0;
    else if (n == "ks_lazy_accumulate")
      ctx->ks_lazy_accumulate = value != 0;
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

int fhe_hmult_relin_rescale(fhe_context *ctx, int count, int moduli_id,
                            const int64_t *a, const int64_t *b,
                            const int64_t *key, int key_moduli_id,
                            int64_t *out) {
  return guarded([&] {
    require(ctx && a && b && key && out, "null argument");
    require(count > 0, "Multiply: empty batch");
    ctx->check_moduli(moduli_id, 3, "Multiply");
    ctx->check_key(moduli_id, key_moduli_id, "Multiply");
    const auto &P = ctx->params;
    const int limbs = P.limbs_at(moduli_id);
    const int ext_limbs = limbs + P.alpha;
    const int64_t N = ctx->n();
    const int64_t duo_l = int64_t(count) * 2 * limbs;
    const int64_t trio_l = int64_t(count) * 3 * limbs;
    const int64_t ext_l =
        int64_t(count) * ctx->ext_slots(moduli_id) * ext_limbs;
    const int64_t acc_l = int64_t(count) * 2 * ext_limbs;
    // trio_c: the coefficient form of the trio; trio keeps the NTT form,
    // whose d2 limbs serve the partitions' own extended limbs
    int64_t *ws = ctx->scratch(
        size_t(2 * duo_l + 2 * trio_l + ext_l + acc_l) * N * 8);
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
        // d0, d1 (lazy) on the side stream once the trio is written; d2
        // (exact, feeds ModUp) here
        fhe::LazySpec lazy_part = d01, exact_part = d01;
        lazy_part.part = 2;
        exact_part.part = 1;
        hip_check(hipEventRecord(ctx->fork_event, ctx->stream), "event");
        hip_check(hipStreamWaitEvent(ctx->side, ctx->fork_event, 0), "event");
        fhe::ntt_inverse(ctx->tables, trio, trio_c, trio_l, limbs, moduli_id,
                         fhe::kInttExit, ctx->side, lazy_part);
        hip_check(hipEventRecord(ctx->join_event, ctx->side), "event");
        joined = false;
        ctx->stage("intt(trio)", [&] {
          fhe::ntt_inverse(ctx->tables, trio, trio_c, trio_l, limbs,
                           moduli_id, fhe::kInttExit, ctx->stream,
                           exact_part);
        });
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
                    3 * limbs * N);
    if (overlap) {
      joined = true;
      hip_check(hipStreamWaitEvent(ctx->stream, ctx->join_event, 0), "event");
    }
    ctx->stage("moddown_relin_rescale", [&] {
      fhe::launch_moddown_relin_rescale(ctx->tables, acc, coef, out, count,
                                        limbs, moduli_id, ctx->stream);
    });
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
    require(count > 0, "Encode: empty batch");
    const auto &P = ctx->params;
    require(level >= 0 && level <= P.max_level,
            "Encode: the target level cannot be reached with the current "
            "engine parameters.");
    const auto e = P.encode_level(level);
    const int mid = P.moduli_id_of_level(level);
    const int limbs = P.limbs_at(mid);
    // the top level of a one-more-level chain leaves limb 0 zero and encodes
    // onto the primes of the level below (src/ckks/core_encode.cpp:64-78)
    const int first = e.top_of_one_more_level ? 1 : 0;
    const int enc_mid =
        e.top_of_one_more_level ? P.moduli_id_of_level(level - 1) : mid;
    require(limbs - first <= fhe::kMaxEwLimbs, "Encode: too many limbs");
    require(P.n % fhe::kBlock == 0, "Encode: N below the kernels' block size");
    const auto &tables = ctx->encode_tables();
    void *work = ctx->aux_scratch(fhe::encode_work_bytes(tables, count));
    ctx->stage("encode", [&] {
      fhe::launch_encode(ctx->tables, tables, values, complex_input ? 1 : 0,
                         count, e.scale_before_tile, limbs, first, enc_mid,
                         e.r_mult_scale_after_tile.data(), out, work,
                         ctx->stream);
    });
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
    const int64_t poly = int64_t(limbs) << P.log_n;
    ctx->stage("elementwise", [&] {
      for (int ct = 0; ct < count; ct++) {
        const int64_t *x = in + ct * polys * poly;
        int64_t *y = out + ct * polys * poly;
        // polynomial 0 takes the scalar, the others are copied
        fhe::launch_elementwise(ctx->tables, fhe::kEwAddScalarReduce, x,
                                nullptr, y, limbs, limbs, mid, c.data(),
                                ctx->stream);
        if (polys > 1 && x != y)
          hip_check(hipMemcpyAsync(y + poly, x + poly,
                                   size_t(polys - 1) * poly * 8,
                                   hipMemcpyDeviceToDevice, ctx->stream),
                    "copy");
      }
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
    // src/ckks/core_encode.cpp:190-210), tiled, entered and transformed
    int64_t *sample = ctx->stash(size_t(N + poly) * 8);
    int64_t *pt = sample + N;
    std::vector<int64_t> host(size_t(N), 0);
    host[size_t(N / 2)] = value;
    hip_check(hipMemcpyAsync(sample, host.data(), size_t(N) * 8,
                             hipMemcpyHostToDevice, ctx->stream),
              "plaintext upload");
    ctx->stage("multiply_imaginary", [&] {
      fhe::launch_tile_unsigned(ctx->tables, sample, pt, 1, limbs, moduli_id,
                                ctx->stream);
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
    // the host vector must outlive the asynchronous copy
    hip_check(hipStreamSynchronize(ctx->stream), "plaintext upload");
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

// ---------------------------------------------------------------------------
// Engine-level compositions of the operations above on single device texts.

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

void ok(int rc) {
  if (rc != FHE_OK)
    throw FheError(rc, fhe_last_error());
}

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

// ---------------------------------------------------------------------------
// Wire format of device data (the reference declares Data::write,
// src/data/data.h:66, but defines no format).
//
//   offset  size  field
//        0     8  magic "FHEAMD\0\1"
//        8     4  version (1)
//       12     4  kind (FHE_WIRE_PLAINTEXT / CIPHERTEXT / KEY)
//       16     8  chain fingerprint: FNV-1a over the context's primes
//       24     4  log2 N
//       28     4  polys
//       32     4  moduli_id of the first limb
//       36     4  limbs per polynomial
//       40     4  level (-1 for keys)
//       44     4  data flags
//       48    16  scale: the x86-64 long double's 10 bytes, zero padded
//       64     8  payload checksum
//       72     -  payload: polys x limbs x N little-endian int64
//
// The checksum is a 64-bit multiply-xorshift fold over the payload words.

namespace {

constexpr uint64_t kWireMagic = 0x0100444d41454846ull; // "FHEAMD\0\1"
constexpr uint32_t kWireVersion = 1;
constexpr size_t kWireHeader = 72;

uint64_t chain_fingerprint(const fhe::CkksParams &P) {
  uint64_t h = 0xcbf29ce484222325ull;
  for (int i = 0; i < P.prime_count; i++) {
    uint64_t v = P.q[i];
    for (int b = 0; b < 8; b++) {
      h ^= (v >> (8 * b)) & 0xff;
      h *= 0x100000001b3ull;
    }
  }
  return h;
}

uint64_t payload_checksum(const int64_t *w, size_t count) {
  uint64_t h = 0x9e3779b97f4a7c15ull ^ count;
  for (size_t i = 0; i < count; i++) {
    h ^= uint64_t(w[i]) + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    h *= 0xbf58476d1ce4e5b9ull;
    h ^= h >> 31;
  }
  return h;
}

template <class T> void put(uint8_t *b, size_t off, T v) {
  std::memcpy(b + off, &v, sizeof(T));
}
template <class T> T get(const uint8_t *b, size_t off) {
  T v;
  std::memcpy(&v, b + off, sizeof(T));
  return v;
}

} // namespace

extern "C" {

int fhe_wire_size(const fhe_context *ctx, int polys, int limbs,
                  size_t *bytes) {
  return guarded([&] {
    require(ctx && bytes && polys > 0 && limbs > 0, "Wire: bad shape");
    *bytes = kWireHeader + size_t(polys) * limbs * size_t(ctx->n()) * 8;
  });
}

int fhe_wire_write(fhe_context *ctx, int kind, const int64_t *data, int polys,
                   int moduli_id, int limbs, int level, int flags,
                   long double scale, void *buffer, size_t capacity,
                   size_t *written) {
  return guarded([&] {
    require(ctx && data && buffer && written, "null argument");
    require(kind >= FHE_WIRE_PLAINTEXT && kind <= FHE_WIRE_KEY,
            "Wire: unknown kind");
    const auto &P = ctx->params;
    require(polys > 0 && limbs > 0 && moduli_id >= 0 &&
                moduli_id < P.prime_count &&
                limbs <= P.prime_count - moduli_id,
            "Wire: shape out of range");
    require(level >= -1 && level <= P.full_level &&
                (flags & ~(FHE_FLAG_NTT_FORM | FHE_FLAG_RESCALED)) == 0,
            "Wire: level or flags out of range");
    const size_t words = size_t(polys) * limbs * size_t(ctx->n());
    const size_t need = kWireHeader + words * 8;
    *written = need;
    require(capacity >= need, "Wire: buffer too small");
    auto *b = static_cast<uint8_t *>(buffer);
    hip_check(hipMemcpyAsync(b + kWireHeader, data, words * 8,
                             hipMemcpyDeviceToHost, ctx->stream),
              "copy to host");
    hip_check(hipStreamSynchronize(ctx->stream), "copy to host");
    std::memset(b, 0, kWireHeader);
    put<uint64_t>(b, 0, kWireMagic);
    put<uint32_t>(b, 8, kWireVersion);
    put<uint32_t>(b, 12, uint32_t(kind));
    put<uint64_t>(b, 16, chain_fingerprint(P));
    put<int32_t>(b, 24, P.log_n);
    put<int32_t>(b, 28, polys);
    put<int32_t>(b, 32, moduli_id);
    put<int32_t>(b, 36, limbs);
    put<int32_t>(b, 40, level);
    put<int32_t>(b, 44, flags);
    std::memcpy(b + 48, &scale, 10);
    put<uint64_t>(b, 64,
                  payload_checksum(
                      reinterpret_cast<const int64_t *>(b + kWireHeader),
                      words));
  });
}

int fhe_wire_read_header(const fhe_context *ctx, const void *buffer,
                         size_t size, fhe_wire_header *header) {
  return guarded([&] {
    require(ctx && buffer && header, "null argument");
    require(size >= kWireHeader, "Wire: truncated header");
    const auto *b = static_cast<const uint8_t *>(buffer);
    require(get<uint64_t>(b, 0) == kWireMagic, "Wire: not an fhe_amd record");
    require(get<uint32_t>(b, 8) == kWireVersion, "Wire: unknown version");
    const auto &P = ctx->params;
    require(get<uint64_t>(b, 16) == chain_fingerprint(P) &&
                get<int32_t>(b, 24) == P.log_n,
            "Wire: the record belongs to another prime chain");
    header->kind = int(get<uint32_t>(b, 12));
    header->polys = get<int32_t>(b, 28);
    header->moduli_id = get<int32_t>(b, 32);
    header->limbs = get<int32_t>(b, 36);
    header->level = get<int32_t>(b, 40);
    header->flags = get<int32_t>(b, 44);
    header->scale = 0;
    std::memcpy(&header->scale, b + 48, 10);
    // every bound is checked without overflow: limbs against the primes
    // left above moduli_id, polys against what the record's size can hold
    require(header->kind >= FHE_WIRE_PLAINTEXT &&
                header->kind <= FHE_WIRE_KEY && header->polys > 0 &&
                header->limbs > 0 && header->moduli_id >= 0 &&
                header->moduli_id < P.prime_count &&
                header->limbs <= P.prime_count - header->moduli_id,
            "Wire: header out of range");
    require(header->level >= -1 && header->level <= P.full_level &&
                (header->flags & ~(FHE_FLAG_NTT_FORM | FHE_FLAG_RESCALED)) ==
                    0,
            "Wire: header out of range");
    const size_t poly_bytes = size_t(header->limbs) * size_t(ctx->n()) * 8;
    require(size_t(header->polys) <= (size - kWireHeader) / poly_bytes,
            "Wire: truncated payload");
    header->payload_bytes = size_t(header->polys) * poly_bytes;
  });
}

int fhe_wire_read(fhe_context *ctx, const void *buffer, size_t size,
                  int64_t *data, size_t capacity) {
  return guarded([&] {
    require(data, "null argument");
    fhe_wire_header h;
    ok(fhe_wire_read_header(ctx, buffer, size, &h));
    require(capacity >= h.payload_bytes, "Wire: device buffer too small");
    const auto *b = static_cast<const uint8_t *>(buffer);
    require(payload_checksum(
                reinterpret_cast<const int64_t *>(b + kWireHeader),
                h.payload_bytes / 8) == get<uint64_t>(b, 64),
            "Wire: payload checksum mismatch");
    hip_check(hipMemcpyAsync(data, b + kWireHeader, h.payload_bytes,
                             hipMemcpyHostToDevice, ctx->stream),
              "copy to device");
    hip_check(hipStreamSynchronize(ctx->stream), "copy to device");
  });
}

} // extern "C"
