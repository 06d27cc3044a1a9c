// Internal helpers shared by the host (C++) and device (HIP) translation units.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/gbo.h"

namespace gbo {

// Records the message of the last failure on this thread and returns `code`.
int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// SplitMix64 finalizer; the generators are counter based so that every
// (seed, stream, index) triple maps to one 64-bit value independent of
// thread count or visiting order.
static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
static inline uint64_t rand64(uint64_t seed, uint64_t stream, uint64_t idx) {
  const uint64_t key = mix64(seed * 0x9e3779b97f4a7c15ULL + stream + 1);
  return mix64(key + (idx + 1) * 0x9e3779b97f4a7c15ULL);
}
// Uniform integer in [0, n) from the high 32 bits (n < 2^32).
static inline uint32_t rand_below(uint64_t r, uint32_t n) {
  return (uint32_t)(((r >> 32) * (uint64_t)n) >> 32);
}

enum : uint64_t {
  STREAM_ER = 1,
  STREAM_RMAT = 2,
  STREAM_ER_M = 3,
  STREAM_WEIGHT = 5,
};

}  // namespace gbo

#define GBO_TRY_BEGIN try {
#define GBO_TRY_END                                                          \
  }                                                                          \
  catch (const std::bad_alloc&) {                                            \
    return gbo::set_error(GBO_ENOMEM, "%s: host allocation failed", __func__); \
  }                                                                          \
  catch (...) {                                                              \
    return gbo::set_error(GBO_EINVAL, "%s: unexpected exception", __func__); \
  }
