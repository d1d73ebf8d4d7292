// The reference-side binding of the MI355X backend: GpuCKKSEngine, a
// CKKSEngine (src/ckks/engine.h) whose multiply(ct, ct, rlk) runs on the
// device through the C ABI (include/fhe_amd.h), compiled against the
// reference's own headers and library by oracle/Makefile (target `binding`)
// the way a maintainer would add src/ckks/engine_gpu.cpp.  It keeps the
// reference's contract: the caller deletes the returned ciphertext, errors
// are std::runtime_error with the library's message.
//
// The device path covers the metric's operation - both inputs rescaled at
// one level, the form the engine's own multiply + rescale pipeline hands on
// (src/ckks/engine_multiply.cpp:134-147 after engine_rescale.cpp:50-60) -
// and every other form is passed to CKKSEngine::multiply.  The relinearization
// key is uploaded once and kept on the device.
//
// gpu_engine_multiply_check (C) runs one multiply(ct, ct, rlk) through the
// binding and through CKKSEngine itself on the same fresh encryptions and
// reports whether limbs and metadata agree; tests/test_gpu_binding.py calls
// it.  Test infrastructure of this repository, like ref_shim.cpp.
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "ckks/engine.h"
#include "engine/engine.h"
#include "fhe_amd.h"

using namespace desilo::fhe;

namespace {

void fhe_check(int code) {
  if (code != FHE_OK)
    throw std::runtime_error(fhe_last_error());
}

class GpuCKKSEngine : public CKKSEngine {
public:
  explicit GpuCKKSEngine(const CKKSConfig &config) : CKKSEngine(config) {
    fhe_params p{};
    p.log_coeff_count = config.log_coeff_count_;
    p.special_prime_count = config.special_prime_count_;
    const bool boot = config.scheme_preset_ == SchemePreset::CKKS_54_60_BOOTSTRAP;
    p.scale_bits = boot ? 54 : 40;
    p.small_prime_bits = boot ? 54 : 40;
    p.one_more_level = 1;
    p.max_level = -1;
    fhe_check(fhe_context_create(&p, &ctx_));
  }
  ~GpuCKKSEngine() override {
    for (auto &kv : keys_)
      (void)fhe_device_free(ctx_, kv.second.data);
    fhe_context_destroy(ctx_);
  }

  using CKKSEngine::multiply;

  Ciphertext *multiply(Ciphertext *a, Ciphertext *b,
                       RelinearizationKey *rlk) override {
    const bool device_form = a->size() == b->size() && a->size() > 0 &&
                             a->flag() == b->flag() &&
                             a->level() == b->level() &&
                             is_rescaled(a->flag()) && !is_ntt_form(a->flag());
    if (!device_form)
      return CKKSEngine::multiply(a, b, rlk);
    std::vector<UnitCiphertext *> units;
    for (int i = 0; i < a->size(); i++)
      units.push_back(multiply_rescaled(a->get(i), b->get(i), rlk));
    return new Ciphertext(units);
  }

  std::string get_mode() const override { return "gpu"; }

  // multiply(ct, ct, rlk) of two rescaled, INTT-form duos at one level on
  // the device: fhe_multiply prepares (NTT), tensors and relinearizes; the
  // result is not rescaled, one level down, at the inputs' moduli_id, with
  // the product of the scales (src/ckks/core.cpp:292-295)
  UnitCiphertext *multiply_rescaled(UnitCiphertext *a, UnitCiphertext *b,
                                    RelinearizationKey *rlk) {
    const int limbs = a->chain_count_;
    const int n = int(a->coeff_count_);
    const size_t duo = size_t(2) * limbs * n;
    int64_t *da = upload(a), *db = upload(b);
    int64_t *dout = alloc(duo);
    const Key &key = resident(rlk);
    const int flags = FHE_FLAG_RESCALED;
    const int rc = fhe_multiply(ctx_, 1, da, a->level_, flags, db, b->level_,
                                flags, key.data, key.moduli_id, dout);
    std::vector<int64_t> host(duo);
    const int rc2 = rc == FHE_OK ? fhe_copy_to_host(ctx_, host.data(), dout,
                                                    duo * 8)
                                 : rc;
    (void)fhe_device_free(ctx_, da);
    (void)fhe_device_free(ctx_, db);
    (void)fhe_device_free(ctx_, dout);
    fhe_check(rc2);
    std::vector<Polynomial *> polys;
    for (int p = 0; p < 2; p++) {
      auto poly = new Polynomial(DataFlag(0), a->moduli_id_, limbs, n);
      for (int c = 0; c < limbs; c++)
        std::memcpy(poly->get(c), host.data() + (size_t(p) * limbs + c) * n,
                    sizeof(int64_t) * n);
      polys.push_back(poly);
    }
    return new UnitCiphertext(polys, DataFlag(0), a->scheme_preset_,
                              a->log_slot_count_, a->log_coeff_count_,
                              a->special_prime_count_, a->moduli_id_,
                              uint8_t(a->level_ - 1), uint8_t(limbs),
                              uint32_t(n), a->ckks_scale_ * b->ckks_scale_);
  }

private:
  struct Key {
    int64_t *data = nullptr;
    int moduli_id = 0;
  };

  int64_t *alloc(size_t words) {
    void *p = nullptr;
    fhe_check(fhe_device_alloc(ctx_, words * 8, &p));
    return static_cast<int64_t *>(p);
  }

  // the limbs of a text as one contiguous [poly][limb][N] device buffer
  int64_t *upload(Data *d) {
    const int polys = int(d->polynomial_count());
    const int limbs = d->chain_count_;
    const size_t n = d->coeff_count_;
    std::vector<int64_t> host(size_t(polys) * limbs * n);
    for (int p = 0; p < polys; p++)
      for (int c = 0; c < limbs; c++)
        std::memcpy(host.data() + (size_t(p) * limbs + c) * n,
                    d->get_polynomial(p)->get(c), n * sizeof(int64_t));
    int64_t *dev = alloc(host.size());
    fhe_check(fhe_copy_to_device(ctx_, dev, host.data(), host.size() * 8));
    return dev;
  }

  const Key &resident(RelinearizationKey *rlk) {
    auto it = keys_.find(rlk);
    if (it != keys_.end())
      return it->second;
    Data *ksk = rlk->key_switching_key_;
    Key k;
    k.data = upload(ksk);
    k.moduli_id = ksk->moduli_id_;
    return keys_[rlk] = k;
  }

  fhe_context *ctx_ = nullptr;
  std::map<RelinearizationKey *, Key> keys_;
};

thread_local std::string g_error;

} // namespace

extern "C" {

const char *gpu_engine_last_error() { return g_error.c_str(); }

// One multiply(ct, ct, rlk) through GpuCKKSEngine and through CKKSEngine on
// the same rescaled encryptions at `level` (N = 2^log_n, alpha special
// primes): returns 0 when limbs, moduli_id, level, flag and scale agree and
// the device result decrypts to x * y within `tolerance`, a positive count
// of differing limbs, or -1 with gpu_engine_last_error() on an error.
// mode_out receives get_mode().
int gpu_engine_multiply_check(int log_n, int alpha, int level,
                              double tolerance, char *mode_out,
                              int mode_capacity) {
  try {
    CKKSConfig config;
    config.log_coeff_count_ = log_n;
    config.special_prime_count_ = alpha;
    GpuCKKSEngine gpu(config);
    std::snprintf(mode_out, size_t(mode_capacity), "%s",
                  gpu.get_mode().c_str());
    auto sk = gpu.create_secret_key();
    auto pk = gpu.create_public_key(sk);
    auto rlk = gpu.create_relinearization_key(sk);
    const int slots = gpu.get_slot_count();
    std::vector<double> x(slots), y(slots);
    for (int i = 0; i < slots; i++) {
      x[i] = std::sin(0.001 * i);
      y[i] = std::cos(0.002 * i);
    }
    Ciphertext *a0 = gpu.encrypt(x, pk, level);
    Ciphertext *b0 = gpu.encrypt(y, pk, level);
    Ciphertext *a = gpu.rescale(a0);
    Ciphertext *b = gpu.rescale(b0);
    Ciphertext *got = gpu.multiply(a, b, rlk);          // device
    Ciphertext *want = gpu.CKKSEngine::multiply(a, b, rlk); // reference
    int diff = 0;
    UnitCiphertext *g = got->get(0), *w = want->get(0);
    if (g->moduli_id_ != w->moduli_id_ || g->level_ != w->level_ ||
        g->flag_ != w->flag_ || g->chain_count_ != w->chain_count_ ||
        g->ckks_scale_ != w->ckks_scale_)
      diff = 1 << 20;
    else
      for (int p = 0; p < 2; p++)
        for (int c = 0; c < g->chain_count_; c++)
          if (std::memcmp(g->get_polynomial(p)->get(c),
                          w->get_polynomial(p)->get(c),
                          sizeof(int64_t) * g->coeff_count_) != 0)
            diff++;
    if (diff == 0) {
      const auto dec = gpu.decrypt(got, sk);
      for (int i = 0; i < slots; i++)
        if (std::abs(dec[i] - x[i] * y[i]) > tolerance) {
          diff = 1 << 21;
          break;
        }
    }
    for (auto t : {a0, b0, a, b, got, want})
      delete t;
    delete rlk;
    delete pk;
    delete sk;
    return diff;
  } catch (const std::exception &e) {
    g_error = e.what();
  }
  return -1;
}

} // extern "C"
