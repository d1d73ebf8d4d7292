#include "context.h"

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
