// Build the SELL-64 form (esr_krylov.h: sell_row) of a row-sorted CSR
// operator on the device: slice widths and offsets (esr_sell_plan), then the
// per-slot column offsets/indices and values (esr_sell_fill), and later the
// values alone when the operator's values have changed (esr_sell_values).
#include "esr_common.h"
#include "esr_krylov.h"

namespace esr {
namespace {

constexpr int kScanThreads = 1024;

// One wavefront per slice: the slice width (longest row), whether every
// column offset col - r fits in int16, whether a row exceeds 255 entries.
// flags[0] = 1: some offset does not fit; flags[1] = 1: a row is too long
// (plain stores of the same value, no read-modify-write).
__global__ __launch_bounds__(256) void k_sell_plan(int64_t n, int64_t ns,
                                                   const int32_t *__restrict__ rowptr,
                                                   const int32_t *__restrict__ col,
                                                   int64_t *__restrict__ width,
                                                   int32_t *__restrict__ flags) {
  const int64_t s = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (s >= ns) return;
  const int64_t r = s * 64 + (threadIdx.x & 63);
  int len = 0;
  bool fits = true;
  if (r < n) {
    const int32_t b = rowptr[r], e = rowptr[r + 1];
    len = e - b;
    for (int32_t j = b; j < e; ++j) {
      const int64_t d = (int64_t)col[j] - r;
      fits = fits && d >= -32768 && d <= 32767;
    }
  }
  int m = len;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(m, off, 64);
    m = o > m ? o : m;
  }
  const bool all_fit = __all(fits);
  if ((threadIdx.x & 63) == 0) {
    width[s] = 64 * (int64_t)m;
    if (!all_fit) flags[0] = 1;
    if (m > 255) flags[1] = 1;
  }
}

// soff[0] = 0, soff[s + 1] = soff[s] + width[s]: one workgroup, each thread
// a contiguous run of slices; in place (width and soff share storage:
// width[s] is read from soff[s + 1] before it is overwritten).
__global__ __launch_bounds__(kScanThreads) void k_sell_scan(int64_t ns, int64_t *__restrict__ soff) {
  __shared__ int64_t sh[kScanThreads];
  const int64_t per = (ns + kScanThreads - 1) / kScanThreads;
  const int64_t lo = threadIdx.x * per;
  const int64_t hi = lo + per < ns ? lo + per : ns;
  int64_t run = 0;
  for (int64_t s = lo; s < hi; ++s) run += soff[s + 1];
  sh[threadIdx.x] = run;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const int64_t v = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += v;
    __syncthreads();
  }
  int64_t acc = threadIdx.x ? sh[threadIdx.x - 1] : 0;
  for (int64_t s = lo; s < hi; ++s) {
    const int64_t w = soff[s + 1];
    acc += w;
    soff[s + 1] = acc;
  }
  if (threadIdx.x == 0) soff[0] = 0;
}

template <typename C>
__device__ __forceinline__ C sell_encode(int64_t r, int64_t c);
template <>
__device__ __forceinline__ int16_t sell_encode<int16_t>(int64_t r, int64_t c) {
  return (int16_t)(c - r);
}
template <>
__device__ __forceinline__ int32_t sell_encode<int32_t>(int64_t r, int64_t c) {
  return (int32_t)c;
}

// Structure and values, one thread per row (rows 64 s .. 64 s + 63 of the
// last slice past n fill their padding with column n - 1).
template <typename T, typename C>
__global__ __launch_bounds__(256) void k_sell_fill(int64_t n, int64_t ns,
                                                   const int32_t *__restrict__ rowptr,
                                                   const int32_t *__restrict__ col,
                                                   const T *__restrict__ Ae,
                                                   const int64_t *__restrict__ soff,
                                                   uint8_t *__restrict__ cnt,
                                                   C *__restrict__ colv, T *__restrict__ val) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t s = r >> 6;
  if (s >= ns) return;
  const int64_t e0 = soff[s] + (r & 63);
  const int w = (int)((soff[s + 1] - soff[s]) >> 6);
  int32_t b = 0, len = 0;
  if (r < n) {
    b = rowptr[r];
    len = rowptr[r + 1] - b;
    cnt[r] = (uint8_t)len;
  }
  const int64_t pad = r < n ? r : n - 1;
  for (int k = 0; k < w; ++k) {
    const int64_t e = e0 + 64 * (int64_t)k;
    if (k < len) {
      colv[e] = sell_encode<C>(r, col[b + k]);
      val[e] = Ae[b + k];
    } else {
      colv[e] = sell_encode<C>(r, pad);
      val[e] = T(0);
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_sell_values(int64_t n, const int32_t *__restrict__ rowptr,
                                                     const T *__restrict__ Ae,
                                                     const int64_t *__restrict__ soff,
                                                     T *__restrict__ val) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  const int64_t e0 = soff[r >> 6] + (r & 63);
  const int32_t b = rowptr[r], len = rowptr[r + 1] - b;
  for (int k = 0; k < len; ++k) val[e0 + 64 * (int64_t)k] = Ae[b + k];
}

} // namespace
} // namespace esr

using namespace esr;

ESR_API int esr_sell_plan(int64_t n, const int32_t *rowptr, const int32_t *col, int64_t *soff,
                          int32_t *flags, int64_t *slots, int32_t *idx16, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1 && rowptr && col && soff && flags && slots && idx16,
              "esr_sell_plan: null argument");
  hipStream_t s = as_stream(stream);
  const int64_t ns = ceil_div(n, 64);
  ESR_HIP(hipMemsetAsync(flags, 0, 2 * sizeof(int32_t), s));
  k_sell_plan<<<(unsigned)ceil_div(ns, 4), 256, 0, s>>>(n, ns, rowptr, col, soff + 1, flags);
  k_sell_scan<<<1, kScanThreads, 0, s>>>(ns, soff);
  ESR_LAUNCHED();
  int32_t hf[2];
  int64_t total = 0;
  ESR_HIP(hipMemcpyAsync(hf, flags, sizeof(hf), hipMemcpyDeviceToHost, s));
  ESR_HIP(hipMemcpyAsync(&total, soff + ns, sizeof(total), hipMemcpyDeviceToHost, s));
  ESR_HIP(hipStreamSynchronize(s));
  if (hf[1]) {
    set_error("esr_sell_plan: a row has more than 255 entries");
    return ESR_ERR_UNSUPPORTED;
  }
  *slots = total;
  *idx16 = hf[0] ? 0 : 1;
  return ESR_OK;
  ESR_GUARD_END
}

ESR_API int esr_sell_fill(int dtype, int64_t n, const int32_t *rowptr, const int32_t *col,
                          const void *Ae, const int64_t *soff, int32_t idx16, uint8_t *cnt,
                          void *colv, void *val, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1 && rowptr && col && Ae && soff && cnt && colv && val,
              "esr_sell_fill: null argument");
  hipStream_t s = as_stream(stream);
  const int64_t ns = ceil_div(n, 64);
  const unsigned g = (unsigned)ceil_div(ns * 64, 256);
#define ESR_SELL_FILL(T)                                                               \
  do {                                                                                 \
    if (idx16)                                                                         \
      k_sell_fill<T, int16_t><<<g, 256, 0, s>>>(n, ns, rowptr, col, (const T *)Ae,    \
                                                 soff, cnt, (int16_t *)colv, (T *)val); \
    else                                                                               \
      k_sell_fill<T, int32_t><<<g, 256, 0, s>>>(n, ns, rowptr, col, (const T *)Ae,    \
                                                 soff, cnt, (int32_t *)colv, (T *)val); \
  } while (0)
  if (dtype == ESR_F64)
    ESR_SELL_FILL(double);
  else if (dtype == ESR_F32)
    ESR_SELL_FILL(float);
  else {
    set_error("esr_sell_fill: float32 or float64 values, got dtype %d", dtype);
    return ESR_ERR_UNSUPPORTED;
  }
#undef ESR_SELL_FILL
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

ESR_API int esr_sell_values(int dtype, int64_t n, const int32_t *rowptr, const void *Ae,
                            const int64_t *soff, void *val, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1 && rowptr && Ae && soff && val, "esr_sell_values: null argument");
  hipStream_t s = as_stream(stream);
  const unsigned g = (unsigned)ceil_div(n, 256);
  if (dtype == ESR_F64)
    k_sell_values<double><<<g, 256, 0, s>>>(n, rowptr, (const double *)Ae, soff, (double *)val);
  else if (dtype == ESR_F32)
    k_sell_values<float><<<g, 256, 0, s>>>(n, rowptr, (const float *)Ae, soff, (float *)val);
  else {
    set_error("esr_sell_values: float32 or float64 values, got dtype %d", dtype);
    return ESR_ERR_UNSUPPORTED;
  }
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}
