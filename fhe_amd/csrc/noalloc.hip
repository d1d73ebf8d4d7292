// An allocation-free form of the batched HMult + relinearize + rescale, for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid allocation).  Its workspace is
// one file-scope device array, `scratch`, of kScratchBytes: every GPU the
// library runs on holds its own copy until the process ends.  Host code
// never uses the array's name as a device address: each device's address
// comes from hipGetSymbolAddress at context creation (with that device
// current) and is kept per device.
//
// Every call on a GPU shares that GPU's array, and the array keeps the last
// call's contents: the pipeline writes every word it reads, so nothing has
// to start at zero.  Calls that overlap on one GPU (two contexts' streams,
// or two host threads) race on the array.
#include <mutex>

#include "context.h"

namespace {

constexpr size_t kScratchBytes = size_t(512) << 20; // 512 MiB
constexpr size_t kScratchWords = kScratchBytes / 8;
constexpr int kMaxDevices = 64;

__device__ __align__(256) int64_t scratch[kScratchWords];

std::mutex scratch_mutex;
int64_t *scratch_address[kMaxDevices] = {};

} // namespace

// The device address of `scratch` on the current device (`device`), looked
// up once per device; called from fhe_context_create.
void noalloc_scratch_setup(int device) {
  require(device >= 0 && device < kMaxDevices,
          "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (scratch_address[device])
    return;
  void *p = nullptr;
  hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
            "hipGetSymbolAddress(scratch)");
  scratch_address[device] = static_cast<int64_t *>(p);
}

extern "C" int fhe_hmult_relin_rescale_noalloc(fhe_context *ctx, int count,
                                               int moduli_id,
                                               const int64_t *a,
                                               const int64_t *b,
                                               const int64_t *key,
                                               int key_moduli_id,
                                               int64_t *out) {
  return guarded([&] {
    require(ctx && a && b && key && out, "null argument");
    require(count > 0, "Multiply: empty batch");
    ctx->check_moduli(moduli_id, 3, "Multiply");
    ctx->check_key(moduli_id, key_moduli_id, "Multiply");
    const int64_t ws = hmult_workspace_words(ctx, count, moduli_id);
    const int64_t digits = ctx->ks_digits_words(count, moduli_id);
    require(size_t(ws + digits) <= kScratchWords,
            "Multiply (no allocation): the batch needs " +
                std::to_string((ws + digits) * 8) +
                " bytes of workspace, above the " +
                std::to_string(kScratchBytes) +
                "-byte scratch; split the batch");
    require(ctx->modup_cache.count({moduli_id, key_moduli_id}) != 0,
            "Multiply (no allocation): no ModUp constants for this key "
            "level (contexts prepare those of the maximum key-switching "
            "level)");
    int64_t *base = nullptr;
    {
      std::lock_guard<std::mutex> lock(scratch_mutex);
      if (ctx->device >= 0 && ctx->device < kMaxDevices)
        base = scratch_address[ctx->device];
    }
    require(base != nullptr, "Multiply (no allocation): no scratch address "
                             "for this device");
    hmult_relin_rescale_into(ctx, count, moduli_id, a, b, key, key_moduli_id,
                             out, base, base + ws);
    ctx->check_launch("Multiply");
  });
}
