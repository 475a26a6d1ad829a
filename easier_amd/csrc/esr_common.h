// Shared helpers of libesrhip.so: error state, launch checks, dtype dispatch,
// XCD-aware block remapping for gfx950.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <exception>
#include <string>

#include <hip/hip_runtime.h>

#include "esr_hip.h"

namespace esr {

// Thread-local message read back through esr_last_error().
void set_error(const char *fmt, ...);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

inline size_t dtype_size(int dtype) {
  switch (dtype) {
  case ESR_F32: return 4;
  case ESR_F64: return 8;
  case ESR_I32: return 4;
  case ESR_I64: return 8;
  case ESR_U8: return 1;
  case ESR_BF16: return 2;
  default: return 0;
  }
}

// MI355X has 8 XCDs; consecutive workgroup ids are dispatched round-robin over
// them. Remapping so each XCD walks one contiguous range of logical tiles keeps
// neighbouring tiles (which share gathered rows) in the same L2.
constexpr int kNumXcd = 8;
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t nb) {
  const int64_t q = nb / kNumXcd, r = nb % kNumXcd;
  const int64_t x = b % kNumXcd, pos = b / kNumXcd;
  const int64_t start = x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q;
  return start + pos;
}

// Workgroup b runs on XCD b % 8 as that XCD's (b / 8)-th: groups of G
// consecutive tiles go to one XCD, XCDs taking consecutive groups, so the 8
// XCDs stream through one window of 8 G tiles together while a tile's
// neighbours mostly stay in its XCD's L2. Tiles past the last whole window
// keep their own index.
__device__ __forceinline__ int64_t group_remap(int64_t b, int64_t nb, int64_t G) {
  const int64_t full = (nb / (kNumXcd * G)) * kNumXcd * G;
  if (b >= full) return b;
  const int64_t x = b % kNumXcd, s = b / kNumXcd;
  return ((s / G) * kNumXcd + x) * G + (s % G);
}

// IEEE correctly rounded square root. The device sqrt can be off by one ulp;
// given y within one ulp of sqrt(x), the Tuckerman rounding test (Markstein,
// "IA-64 and Elementary Functions", 2000) picks the correctly rounded
// neighbour with two exact-sign FMA residuals: round up when x > y * y_up,
// down when x <= y * y_down (the half-ulp-squared terms of the exact test
// are below the granularity of x - y * y_neighbour).
// (x restricted to a range where y is a positive normal number, so the
// neighbours are the bit patterns one above and one below.)
__device__ __forceinline__ double cr_sqrt(double x) {
  const double y = sqrt(x);
  if (!(x > 0x1p-900) || x > 0x1p+1000) return y;
  const long long yb = __builtin_bit_cast(long long, y);
  const double yu = __builtin_bit_cast(double, yb + 1);
  const double yd = __builtin_bit_cast(double, yb - 1);
  if (__builtin_fma(-y, yu, x) > 0.0) return yu;
  if (__builtin_fma(-y, yd, x) <= 0.0) return yd;
  return y;
}
__device__ __forceinline__ float cr_sqrt(float x) {
  const float y = sqrtf(x);
  if (!(x > 0x1p-100f) || x > 0x1p+100f) return y;
  const int yb = __builtin_bit_cast(int, y);
  const float yu = __builtin_bit_cast(float, yb + 1);
  const float yd = __builtin_bit_cast(float, yb - 1);
  if (__builtin_fmaf(-y, yu, x) > 0.0f) return yu;
  if (__builtin_fmaf(-y, yd, x) <= 0.0f) return yd;
  return y;
}

} // namespace esr

#define ESR_API extern "C" __attribute__((visibility("default")))

#define ESR_REQUIRE(cond, ...)                                                 \
  do {                                                                         \
    if (!(cond)) {                                                             \
      esr::set_error(__VA_ARGS__);                                             \
      return ESR_ERR_INVALID;                                                  \
    }                                                                          \
  } while (0)

#define ESR_HIP(expr)                                                          \
  do {                                                                         \
    hipError_t _e = (expr);                                                    \
    if (_e != hipSuccess) {                                                    \
      esr::set_error("%s: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, \
                     __LINE__);                                                \
      return ESR_ERR_HIP;                                                      \
    }                                                                          \
  } while (0)

#define ESR_LAUNCHED() ESR_HIP(hipGetLastError())

// Wraps an ABI body so no C++ exception escapes.
#define ESR_GUARD_BEGIN try {
#define ESR_GUARD_END                                                          \
  }                                                                            \
  catch (const std::exception &ex) {                                           \
    esr::set_error("internal error: %s", ex.what());                           \
    return ESR_ERR_INTERNAL;                                                   \
  }                                                                            \
  catch (...) {                                                                \
    esr::set_error("internal error");                                          \
    return ESR_ERR_INTERNAL;                                                   \
  }
