// Shared host/device helpers for libsmx: error reporting and launch checks.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdarg>
#include <cstdio>
#include <cstdint>

namespace smx {

void set_error(const char* fmt, ...);

#define SMX_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      ::smx::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,                  \
                       hipGetErrorString(e_));                                     \
      return 1;                                                                    \
    }                                                                              \
  } while (0)

#define SMX_REQUIRE(cond, ...)                                                     \
  do {                                                                             \
    if (!(cond)) {                                                                 \
      ::smx::set_error(__VA_ARGS__);                                               \
      return 2;                                                                    \
    }                                                                              \
  } while (0)

#define SMX_LAUNCHED()                                                             \
  do {                                                                             \
    hipError_t e_ = hipGetLastError();                                             \
    if (e_ != hipSuccess) {                                                        \
      ::smx::set_error("%s:%d launch: %s", __FILE__, __LINE__,                     \
                       hipGetErrorString(e_));                                     \
      return 3;                                                                    \
    }                                                                              \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned blocks_for(int64_t n, int threads) {
  return static_cast<unsigned>((n + threads - 1) / threads);
}

// NaN-propagating max/min, as jnp.maximum / jnp.minimum behave.
__device__ __forceinline__ double nmax(double a, double b) {
  return (a != a) ? a : ((b != b) ? b : (a > b ? a : b));
}
__device__ __forceinline__ double nmin(double a, double b) {
  return (a != a) ? a : ((b != b) ? b : (a < b ? a : b));
}

// Bijection of [0, G): block b, dispatched to XCD b % 8, gets the index that
// makes each XCD's blocks one contiguous run (the first G % 8 XCDs take
// G / 8 + 1 blocks, the rest G / 8).
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t G) {
  const int64_t q = G / 8, r = G % 8, x = b % 8, slot = b / 8;
  return x < r ? x * (q + 1) + slot : r * (q + 1) + (x - r) * q + slot;
}

}  // namespace smx
