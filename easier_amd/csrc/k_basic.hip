// Primitive kernels behind esr.Selector / esr.Reducer / aggregators and the
// elementwise operators of a traced forward.
//
// Layout conventions: every tensor is row-major (n, width) with the mesh index
// as the leading dimension; indices are int32. Every launch below is one tile
// per workgroup (no grid-stride loops) so the XCD remap places consecutive
// tiles on one XCD and its L2.
#include <cmath>
#include <limits>
#include <mutex>
#include <type_traits>

#include "esr_common.h"

namespace esr {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

// ----------------------------------------------------------------------------
// wave / block reductions with a fixed combination tree (deterministic)
// ----------------------------------------------------------------------------
template <typename T, typename Op>
__device__ __forceinline__ T wave_reduce(T v, Op op) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_down(v, off, 64));
  return v;
}

// Result valid in thread 0 only.
template <typename T, typename Op>
__device__ __forceinline__ T block_reduce(T v, Op op, T identity) {
  __shared__ T part[kWaves];
  v = wave_reduce(v, op);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) part[wid] = v;
  __syncthreads();
  T r = identity;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < kWaves; ++w) r = op(r, part[w]);
  }
  return r;
}

// ----------------------------------------------------------------------------
// a1: gather
// ----------------------------------------------------------------------------
// One row per element (width*elem bytes == sizeof(U)); 4 independent gathers
// in flight per thread, each instruction coalesced over the wave.
template <typename U>
__global__ __launch_bounds__(kBlock) void k_gather_row(const U *__restrict__ src,
                                                      const int32_t *__restrict__ idx,
                                                      U *__restrict__ dst, int64_t n) {
  const int64_t tile = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t base = tile * (4 * kBlock) + threadIdx.x;
  int32_t i[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t e = base + k * kBlock;
    i[k] = e < n ? __builtin_nontemporal_load(idx + e) : 0;
  }
  U v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = (base + k * kBlock < n) ? src[i[k]] : U{};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t e = base + k * kBlock;
    if (e < n) dst[e] = v[k];
  }
}

// Rows of `upr` units.
template <typename U>
__global__ __launch_bounds__(kBlock) void k_gather_units(const U *__restrict__ src,
                                                        const int32_t *__restrict__ idx,
                                                        U *__restrict__ dst, int64_t n,
                                                        int64_t upr) {
  const int64_t tile = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t total = n * upr;
  const int64_t base = tile * (4 * kBlock) + threadIdx.x;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t t = base + k * kBlock;
    if (t < total) {
      const int64_t e = t / upr, c = t - e * upr;
      dst[t] = src[(int64_t)idx[e] * upr + c];
    }
  }
}

// One-unit rows, V consecutive outputs per vector (V units = 16 bytes): a
// lane loads V indices with one load, gathers V rows and writes them with
// one 16-byte store; two vectors per lane, kBlock vectors apart. The n % V
// tail is written by the lane holding vector index n / V.
typedef int32_t i32x2_t __attribute__((ext_vector_type(2)));
typedef int32_t i32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
template <int V> struct GatherIdx;
template <> struct GatherIdx<2> { typedef i32x2_t I; };
template <> struct GatherIdx<4> { typedef i32x4_t I; };

template <typename U, int V>
__global__ __launch_bounds__(kBlock) void k_gather_vec(const U *__restrict__ src,
                                                      const int32_t *__restrict__ idx,
                                                      U *__restrict__ dst, int64_t n) {
  static_assert(sizeof(U) * V == 16, "16-byte output vectors");
  typedef typename GatherIdx<V>::I IV;
  // groups of 64 consecutive tiles per XCD: the gathered neighbours of a
  // tile mostly sit in its XCD's L2, and the 8 XCDs still stream through one
  // window together. Measured on the 25.2 M-edge float64 gather: 57.2 us;
  // one contiguous range per XCD (xcd_remap) 59.4 us; launch order 63.2 us.
  const int64_t tile = group_remap(blockIdx.x, gridDim.x, 64);
  const int64_t nvec = n / V;
  int64_t vi[2];
  IV ii[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    vi[k] = tile * (2 * kBlock) + k * kBlock + threadIdx.x;
    if (vi[k] < nvec) ii[k] = __builtin_nontemporal_load(reinterpret_cast<const IV *>(idx) + vi[k]);
  }
  U v[2][V];
#pragma unroll
  for (int k = 0; k < 2; ++k)
    if (vi[k] < nvec) {
#pragma unroll
      for (int c = 0; c < V; ++c) v[k][c] = src[ii[k][c]];
    }
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    if (vi[k] < nvec) {
      u32x4_t w;
      __builtin_memcpy(&w, v[k], 16);
      __builtin_nontemporal_store(w, reinterpret_cast<u32x4_t *>(dst) + vi[k]);
    } else if (vi[k] == nvec) {
      for (int64_t e = nvec * V; e < n; ++e) dst[e] = src[idx[e]];
    }
  }
}

template <typename U>
static int launch_gather(const void *src, const int32_t *idx, void *dst,
                         int64_t n_out, int64_t upr, hipStream_t s) {
  const int64_t total = n_out * upr;
  const int64_t grid = ceil_div(total, 4 * kBlock);
  constexpr int V = 16 / (int)sizeof(U);
  const bool vec = upr == 1 && (V == 2 || V == 4) &&
                   (uintptr_t)dst % 16 == 0 && (uintptr_t)idx % (4 * V) == 0;
  if (vec) {
    if constexpr (V == 2 || V == 4) {
      const int64_t vgrid = ceil_div(n_out / V + 1, 2 * kBlock);
      k_gather_vec<U, V><<<dim3((unsigned)vgrid), dim3(kBlock), 0, s>>>(
          (const U *)src, idx, (U *)dst, n_out);
    }
  } else if (upr == 1)
    k_gather_row<U><<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>(
        (const U *)src, idx, (U *)dst, n_out);
  else
    k_gather_units<U><<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>(
        (const U *)src, idx, (U *)dst, n_out, upr);
  ESR_LAUNCHED();
  return ESR_OK;
}

// ----------------------------------------------------------------------------
// a2: CSR reduce
// ----------------------------------------------------------------------------
template <typename T> struct Lim {
  __device__ static T lowest() { return std::numeric_limits<T>::lowest(); }
  __device__ static T highest() { return std::numeric_limits<T>::max(); }
};
template <> struct Lim<float> {
  __device__ static float lowest() { return -INFINITY; }
  __device__ static float highest() { return INFINITY; }
};
template <> struct Lim<double> {
  __device__ static double lowest() { return -(double)INFINITY; }
  __device__ static double highest() { return (double)INFINITY; }
};

template <typename T> __device__ __forceinline__ bool is_nan(T v) {
  if constexpr (std::is_floating_point<T>::value) return v != v;
  else return false;
}

template <typename T, int OP> struct RedOp {
  __device__ static T identity() {
    if constexpr (OP == ESR_RED_PROD) return T(1);
    else if constexpr (OP == ESR_RED_AMAX) return Lim<T>::lowest();
    else if constexpr (OP == ESR_RED_AMIN) return Lim<T>::highest();
    else return T(0);
  }
  // NaN-propagating like torch's scatter_reduce amax/amin.
  __device__ T operator()(T a, T b) const {
    if constexpr (OP == ESR_RED_PROD) return a * b;
    else if constexpr (OP == ESR_RED_AMAX) return (is_nan(b) || b > a) ? b : a;
    else if constexpr (OP == ESR_RED_AMIN) return (is_nan(b) || b < a) ? b : a;
    else return a + b;
  }
};

template <typename T, int OP>
__global__ __launch_bounds__(kBlock) void k_csr_seq(const T *__restrict__ vals, int64_t width,
                                                   const int32_t *__restrict__ rowptr,
                                                   const int32_t *__restrict__ perm,
                                                   int64_t n_rows, int keep_empty,
                                                   T *__restrict__ out) {
  const int64_t tile = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t t = tile * kBlock + threadIdx.x;
  if (t >= n_rows * width) return;
  const int64_t r = width == 1 ? t : t / width;
  const int64_t c = t - r * width;
  const int32_t b = rowptr[r], e = rowptr[r + 1];
  if (e - b > ESR_SEQ_ROW_MAX) return; // handled by k_csr_long
  if (b == e) {
    if (!keep_empty) out[t] = T(0);
    return;
  }
  RedOp<T, OP> op;
  T acc = RedOp<T, OP>::identity();
  if (perm) {
    for (int32_t j = b; j < e; ++j) acc = op(acc, vals[(int64_t)perm[j] * width + c]);
  } else {
    for (int32_t j = b; j < e; ++j) acc = op(acc, vals[(int64_t)j * width + c]);
  }
  if constexpr (OP == ESR_RED_MEAN) acc = acc / T(e - b);
  out[t] = acc;
}

// One-value rows without a permutation (the compiled Reducer's usual form):
// each wavefront owns 64 consecutive rows, loads their whole entry segment
// [rowptr[r0], rowptr[r0 + 64]) with contiguous (coalesced) loads into LDS,
// then every lane folds its own row from LDS in entry order -- the same
// sequential sum as k_csr_seq, with a third of the address work for rows of
// ~3 entries. Segments longer than kCap entries fall back to direct loads.
// (128 rows per wavefront, two per lane, measured slower: 59.8 vs 54.4 us
// for the 8.4 M-row, 25.2 M-entry float64 sum.)
template <typename T, int OP, bool VEC>
__global__ __launch_bounds__(kBlock) void k_csr_staged(const T *__restrict__ vals,
                                                      const int32_t *__restrict__ rowptr,
                                                      int64_t n_rows, int keep_empty,
                                                      T *__restrict__ out) {
  constexpr int kCap = 512, kU = 4;
  constexpr int V = VEC ? 16 / (int)sizeof(T) : 1;   // values per staged load
  typedef T tv_t __attribute__((ext_vector_type(V)));
  __shared__ T stage[kBlock / 64][kCap];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // Tiles in launch order: the segment sum reuses nothing across tiles, and
  // the 8 XCDs streaming through one region together beat one contiguous
  // range per XCD (xcd_remap): 51.9 vs 54.0 us for the 8.4 M-row, 25.2 M-entry
  // float64 sum (groups of 16 or 64 tiles per XCD: 52.1 us).
  const int64_t tile = blockIdx.x;
  const int64_t r0 = (tile * (kBlock / 64) + w) * 64;
  if (r0 >= n_rows) return;                       // the whole wavefront
  const int64_t r = r0 + lane;
  const bool ok = r < n_rows;
  const int64_t rend = r0 + 64 < n_rows ? r0 + 64 : n_rows;
  const int32_t s0 = rowptr[r0], s1 = rowptr[rend];
  const int32_t b = ok ? rowptr[r] : 0;
  const int32_t bn = __shfl_down(b, 1, 64);       // rowptr[r + 1] of the next lane
  const int32_t e = ok ? (r + 1 < rend ? bn : s1) : 0;
  // staged entries [a0, a0 + len): a0 = s0 rounded down to a 16-byte boundary
  const int32_t a0 = s0 & ~(V - 1);
  const int32_t nv = (s1 - a0 + V - 1) / V, len = nv * V;
  T *st = stage[w];
  const bool staged = len <= kCap;                // uniform over the wavefront
  if (staged) {
    const int32_t nnz = rowptr[n_rows];
    auto load = [&](int32_t k) -> tv_t {           // vector k: entries a0 + kV ..
      const int64_t j = (int64_t)a0 + (int64_t)k * V;
      if (j + V <= nnz) return *reinterpret_cast<const tv_t *>(vals + j);
      tv_t x;
#pragma unroll
      for (int c = 0; c < V; ++c) x[c] = j + c < nnz ? vals[j + c] : T(0);
      return x;
    };
    auto put = [&](int32_t k, const tv_t &x) {
#pragma unroll
      for (int c = 0; c < V; ++c) st[k * V + c] = x[c];
    };
    // kU loads per lane in flight together, then the rest
    tv_t reg[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int32_t k = lane + 64 * u;
      if (k < nv) reg[u] = load(k);
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int32_t k = lane + 64 * u;
      if (k < nv) put(k, reg[u]);
    }
    for (int32_t k = lane + 64 * kU; k < nv; k += 64) put(k, load(k));
    // LDS written and read by lanes of one wavefront: order the accesses
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  if (!ok || e - b > ESR_SEQ_ROW_MAX) return;     // long rows: k_csr_long
  if (b == e) {
    if (!keep_empty) out[r] = T(0);
    return;
  }
  RedOp<T, OP> op;
  T acc = RedOp<T, OP>::identity();
  if (staged) {
    for (int32_t j = b; j < e; ++j) acc = op(acc, st[j - a0]);
  } else {
    for (int32_t j = b; j < e; ++j) acc = op(acc, vals[j]);
  }
  if constexpr (OP == ESR_RED_MEAN) acc = acc / T(e - b);
  __builtin_nontemporal_store(acc, out + r);
}

template <typename T, int OP>
__global__ __launch_bounds__(kBlock) void k_csr_long(const T *__restrict__ vals, int64_t width,
                                                    const int32_t *__restrict__ rowptr,
                                                    const int32_t *__restrict__ perm,
                                                    const int32_t *__restrict__ long_rows,
                                                    T *__restrict__ out) {
  const int32_t r = long_rows[blockIdx.x];
  const int64_t c = blockIdx.y;
  const int32_t b = rowptr[r], e = rowptr[r + 1];
  RedOp<T, OP> op;
  T acc = RedOp<T, OP>::identity();
  for (int32_t j = b + threadIdx.x; j < e; j += kBlock) {
    const int64_t src = perm ? (int64_t)perm[j] : (int64_t)j;
    acc = op(acc, vals[src * width + c]);
  }
  acc = block_reduce(acc, op, RedOp<T, OP>::identity());
  if (threadIdx.x == 0) {
    if constexpr (OP == ESR_RED_MEAN) acc = acc / T(e - b);
    out[(int64_t)r * width + c] = acc;
  }
}

template <typename T, int OP>
static int launch_csr(const void *vals, int64_t width, const int32_t *rowptr,
                      const int32_t *perm, int64_t n_rows, const int32_t *long_rows,
                      int64_t n_long, int keep_empty, void *out, hipStream_t s) {
  const int64_t total = n_rows * width;
  if (total > 0 && width == 1 && !perm) {
    const dim3 grid((unsigned)ceil_div(n_rows, kBlock));
    if ((uintptr_t)vals % 16 == 0 && (sizeof(T) == 4 || sizeof(T) == 8))
      k_csr_staged<T, OP, true><<<grid, dim3(kBlock), 0, s>>>((const T *)vals, rowptr, n_rows,
                                                             keep_empty, (T *)out);
    else
      k_csr_staged<T, OP, false><<<grid, dim3(kBlock), 0, s>>>((const T *)vals, rowptr, n_rows,
                                                              keep_empty, (T *)out);
    ESR_LAUNCHED();
  } else if (total > 0) {
    k_csr_seq<T, OP><<<dim3((unsigned)ceil_div(total, kBlock)), dim3(kBlock), 0, s>>>(
        (const T *)vals, width, rowptr, perm, n_rows, keep_empty, (T *)out);
    ESR_LAUNCHED();
  }
  if (n_long > 0) {
    k_csr_long<T, OP><<<dim3((unsigned)n_long, (unsigned)width), dim3(kBlock), 0, s>>>(
        (const T *)vals, width, rowptr, perm, long_rows, (T *)out);
    ESR_LAUNCHED();
  }
  return ESR_OK;
}

template <typename T>
static int dispatch_csr_op(int op, const void *vals, int64_t width, const int32_t *rowptr,
                           const int32_t *perm, int64_t n_rows, const int32_t *long_rows,
                           int64_t n_long, int keep_empty, void *out, hipStream_t s) {
  switch (op) {
  case ESR_RED_SUM:
    return launch_csr<T, ESR_RED_SUM>(vals, width, rowptr, perm, n_rows, long_rows, n_long,
                                      keep_empty, out, s);
  case ESR_RED_PROD:
    return launch_csr<T, ESR_RED_PROD>(vals, width, rowptr, perm, n_rows, long_rows, n_long,
                                       keep_empty, out, s);
  case ESR_RED_MEAN:
    if constexpr (std::is_floating_point<T>::value)
      return launch_csr<T, ESR_RED_MEAN>(vals, width, rowptr, perm, n_rows, long_rows,
                                         n_long, keep_empty, out, s);
    set_error("esr_csr_reduce: 'mean' needs a floating-point dtype");
    return ESR_ERR_UNSUPPORTED;
  case ESR_RED_AMAX:
    return launch_csr<T, ESR_RED_AMAX>(vals, width, rowptr, perm, n_rows, long_rows, n_long,
                                       keep_empty, out, s);
  case ESR_RED_AMIN:
    return launch_csr<T, ESR_RED_AMIN>(vals, width, rowptr, perm, n_rows, long_rows, n_long,
                                       keep_empty, out, s);
  default:
    set_error("esr_csr_reduce: unknown op %d", op);
    return ESR_ERR_INVALID;
  }
}

// ----------------------------------------------------------------------------
// a4: aggregators (two fixed-shape stages)
// ----------------------------------------------------------------------------
constexpr int64_t kAggMaxBlocks = 1024;
constexpr int64_t kAggRowsPerBlock = 8 * kBlock;

static int64_t agg_blocks(int64_t n) {
  int64_t nb = ceil_div(n, kAggRowsPerBlock);
  if (nb < 1) nb = 1;
  if (nb > kAggMaxBlocks) nb = kAggMaxBlocks;
  return nb;
}

// Stage-1 transform of one element.
template <typename T>
__device__ __forceinline__ T agg_pre(int op, double p, T v) {
  if (op != ESR_AGG_NORM) return v;
  const T a = v < T(0) ? -v : v;
  if (p == 2.0) return v * v;
  if (p == 1.0) return a;
  if (isinf(p)) return a;
  if (p == 0.0) return a != T(0) ? T(1) : T(0);
  return (T)pow((double)a, p);
}

template <typename T> struct AggCombine {
  int op;
  double p;
  __device__ T operator()(T a, T b) const {
    switch (op) {
    case ESR_AGG_PROD: return a * b;
    case ESR_AGG_MAX: return (is_nan(b) || b > a) ? b : a;
    case ESR_AGG_MIN: return (is_nan(b) || b < a) ? b : a;
    case ESR_AGG_NORM:
      if (isinf(p)) {
        if (p > 0) return (is_nan(b) || b > a) ? b : a;
        return (is_nan(b) || b < a) ? b : a;
      }
      return a + b;
    default: return a + b;
    }
  }
  __device__ T identity() const {
    switch (op) {
    case ESR_AGG_PROD: return T(1);
    case ESR_AGG_MAX: return Lim<T>::lowest();
    case ESR_AGG_MIN: return Lim<T>::highest();
    case ESR_AGG_NORM:
      if (isinf(p)) return p > 0 ? T(0) : Lim<T>::highest();
      return T(0);
    default: return T(0);
    }
  }
};

template <typename T>
__global__ __launch_bounds__(kBlock) void k_agg_stage1(const T *__restrict__ x, int64_t n,
                                                      int64_t width, int op, double p,
                                                      T *__restrict__ part) {
  const int64_t nb = gridDim.x, b = blockIdx.x, c = blockIdx.y;
  const int64_t chunk = (n + nb - 1) / nb;
  const int64_t lo = b * chunk, hi = lo + chunk < n ? lo + chunk : n;
  AggCombine<T> cmb{op, p};
  T acc = cmb.identity();
  // four loads in flight per step, folded in the same per-thread order as
  // the plain strided loop (i, i + kBlock, ...)
  int64_t i = lo + threadIdx.x;
  for (; i + 3 * kBlock < hi; i += 4 * kBlock) {
    const T v0 = x[i * width + c], v1 = x[(i + kBlock) * width + c],
            v2 = x[(i + 2 * kBlock) * width + c], v3 = x[(i + 3 * kBlock) * width + c];
    acc = cmb(acc, agg_pre<T>(op, p, v0));
    acc = cmb(acc, agg_pre<T>(op, p, v1));
    acc = cmb(acc, agg_pre<T>(op, p, v2));
    acc = cmb(acc, agg_pre<T>(op, p, v3));
  }
  for (; i < hi; i += kBlock) acc = cmb(acc, agg_pre<T>(op, p, x[i * width + c]));
  acc = block_reduce(acc, cmb, cmb.identity());
  if (threadIdx.x == 0) part[c * nb + b] = acc;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_agg_stage2(const T *__restrict__ part, int64_t nb,
                                                      int op, double p, T *__restrict__ out) {
  const int64_t c = blockIdx.x;
  AggCombine<T> cmb{op, p};
  T acc = cmb.identity();
  for (int64_t i = threadIdx.x; i < nb; i += kBlock) acc = cmb(acc, part[c * nb + i]);
  acc = block_reduce(acc, cmb, cmb.identity());
  if (threadIdx.x == 0) {
    if (op == ESR_AGG_NORM && !isinf(p)) {
      if (p == 2.0) acc = (T)cr_sqrt((double)acc);
      else if (p != 1.0 && p != 0.0) acc = (T)pow((double)acc, 1.0 / p);
    }
    out[c] = acc;
  }
}

// ----------------------------------------------------------------------------
// elementwise
// ----------------------------------------------------------------------------
struct EwParams {
  int ndim;
  int op;
  int n_in;
  int64_t total;
  int64_t shape[ESR_EW_MAXDIM];
  int64_t ostr[ESR_EW_MAXDIM];
  int64_t istr[3][ESR_EW_MAXDIM];
  const void *in[3];
  double sc[3];
  int contiguous; // out and tensor inputs dense 1-d (stride 1) or broadcast (stride 0)
};

template <typename T> __device__ __forceinline__ T ew_unary(int op, T a) {
  if constexpr (std::is_floating_point<T>::value) {
    switch (op) {
    case ESR_EW_NEG: return -a;
    case ESR_EW_ABS: return a < T(0) ? -a : (a == T(0) ? T(0) : a);
    case ESR_EW_EXP: return exp(a);
    case ESR_EW_LOG: return log(a);
    case ESR_EW_SQRT: return cr_sqrt(a);
    case ESR_EW_RSQRT: return T(1) / cr_sqrt(a);
    case ESR_EW_SIN: return sin(a);
    case ESR_EW_COS: return cos(a);
    case ESR_EW_TANH: return tanh(a);
    case ESR_EW_SIGN: return a > T(0) ? T(1) : (a < T(0) ? T(-1) : (a != a ? a : T(0)));
    case ESR_EW_RECIP: return T(1) / a;
    case ESR_EW_FLOOR: return floor(a);
    case ESR_EW_CEIL: return ceil(a);
    case ESR_EW_SIGMOID: return T(1) / (T(1) + exp(-a));
    case ESR_EW_RELU: return a > T(0) ? a : (a != a ? a : T(0));
    default: return a;
    }
  } else {
    switch (op) {
    case ESR_EW_NEG: return -a;
    case ESR_EW_ABS: return a < T(0) ? -a : a;
    case ESR_EW_SIGN: return a > T(0) ? T(1) : (a < T(0) ? T(-1) : T(0));
    case ESR_EW_RELU: return a > T(0) ? a : T(0);
    default: return a;
    }
  }
}

template <typename T> __device__ __forceinline__ T ew_pow(T a, T b) {
  if constexpr (std::is_floating_point<T>::value) {
    // Same special cases as the CPU kernels the reference runs on
    // (an exponent of 2 or 3 is evaluated by multiplication).
    if (b == T(2)) return a * a;
    if (b == T(3)) return a * a * a;
    if (b == T(0.5)) return cr_sqrt(a);
    if (b == T(-0.5)) return T(1) / cr_sqrt(a);
    if (b == T(-1)) return T(1) / a;
    if (b == T(-2)) return T(1) / (a * a);
    return pow(a, b);
  } else {
    // torch's integer power: a negative exponent gives 1 for a = 1, +-1 for
    // a = -1 and 0 otherwise; else exponentiation by squaring (at most 64
    // steps), wrapping modulo 2^bits like torch's integer arithmetic
    if (b < T(0)) {
      if (a == T(1)) return T(1);
      if (a == T(-1)) return (b & T(1)) ? T(-1) : T(1);
      return T(0);
    }
    using U = typename std::make_unsigned<T>::type;
    U base = (U)a, r = 1, e = (U)b;
    while (e) {
      if (e & 1u) r *= base;
      base *= base;
      e >>= 1;
    }
    return (T)r;
  }
}

template <typename T> __device__ __forceinline__ T ew_binary(int op, T a, T b) {
  switch (op) {
  case ESR_EW_ADD: return a + b;
  case ESR_EW_SUB: return a - b;
  case ESR_EW_MUL: return a * b;
  case ESR_EW_DIV: return a / b;
  case ESR_EW_POW: return ew_pow(a, b);
  case ESR_EW_MAXIMUM: return (is_nan(a) || a > b) ? a : (is_nan(b) ? b : (a == b ? a : b));
  case ESR_EW_MINIMUM: return (is_nan(a) || a < b) ? a : (is_nan(b) ? b : (a == b ? a : b));
  case ESR_EW_ATAN2:
    if constexpr (std::is_floating_point<T>::value) return atan2(a, b);
    else return a;
  case ESR_EW_FMOD:
    if constexpr (std::is_floating_point<T>::value) return fmod(a, b);
    else return a % b;
  case ESR_EW_FLOORDIV:
    if constexpr (std::is_floating_point<T>::value) return floor(a / b);
    else {
      T q = a / b;
      if ((a % b != 0) && ((a < 0) != (b < 0))) q -= 1;
      return q;
    }
  default: return a;
  }
}

template <typename T> __device__ __forceinline__ bool ew_pred(int op, T a, T b) {
  switch (op) {
  case ESR_EW_LT: return a < b;
  case ESR_EW_LE: return a <= b;
  case ESR_EW_GT: return a > b;
  case ESR_EW_GE: return a >= b;
  case ESR_EW_EQ: return a == b;
  case ESR_EW_NE: return a != b;
  case ESR_EW_AND: return (a != T(0)) && (b != T(0));
  case ESR_EW_OR: return (a != T(0)) || (b != T(0));
  default: return false;
  }
}

__device__ __forceinline__ bool ew_is_pred(int op) {
  return (op >= ESR_EW_LT && op <= ESR_EW_OR) || op == ESR_EW_ISNAN || op == ESR_EW_NOT;
}

template <typename TI, typename TO>
__device__ __forceinline__ TO ew_cast(TI v) {
  if constexpr (std::is_same<TO, uint8_t>::value) return v != TI(0) ? 1 : 0;
  else return (TO)v;
}

template <typename TI, typename TO>
__global__ __launch_bounds__(kBlock) void k_ewise(EwParams p, TO *__restrict__ out) {
  const int64_t tile = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t t = tile * kBlock + threadIdx.x;
  if (t >= p.total) return;
  int64_t off[4] = {0, 0, 0, 0};
  if (p.contiguous) {
    off[0] = t;
#pragma unroll
    for (int k = 0; k < 3; ++k) off[k + 1] = p.istr[k][0] ? t : 0;
  } else {
    int64_t rem = t;
    for (int d = p.ndim - 1; d >= 0; --d) {
      const int64_t sz = p.shape[d];
      const int64_t i = rem % sz;
      rem /= sz;
      off[0] += i * p.ostr[d];
#pragma unroll
      for (int k = 0; k < 3; ++k) off[k + 1] += i * p.istr[k][d];
    }
  }
  const int op = p.op;
  if (op == ESR_EW_WHERE) {
    const uint8_t cnd = p.in[0] ? ((const uint8_t *)p.in[0])[off[1]] : (uint8_t)(p.sc[0] != 0.0);
    const TI a = p.in[1] ? ((const TI *)p.in[1])[off[2]] : (TI)p.sc[1];
    const TI b = p.in[2] ? ((const TI *)p.in[2])[off[3]] : (TI)p.sc[2];
    out[off[0]] = (TO)(cnd ? a : b);
    return;
  }
  const TI a = p.in[0] ? ((const TI *)p.in[0])[off[1]] : (TI)p.sc[0];
  if (op < ESR_EW_ADD) {
    if (op == ESR_EW_COPY) {
      out[off[0]] = ew_cast<TI, TO>(a);
    } else if (op == ESR_EW_ISNAN) {
      out[off[0]] = (TO)(is_nan(a) ? 1 : 0);
    } else if (op == ESR_EW_NOT) {
      out[off[0]] = (TO)(a == TI(0) ? 1 : 0);
    } else {
      out[off[0]] = (TO)ew_unary<TI>(op, a);
    }
    return;
  }
  const TI b = p.in[1] ? ((const TI *)p.in[1])[off[2]] : (TI)p.sc[1];
  if (ew_is_pred(op)) out[off[0]] = (TO)(ew_pred<TI>(op, a, b) ? 1 : 0);
  else out[off[0]] = (TO)ew_binary<TI>(op, a, b);
}

template <typename TI, typename TO>
static int launch_ew(const EwParams &p, void *out, hipStream_t s) {
  if (p.total == 0) return ESR_OK;
  k_ewise<TI, TO><<<dim3((unsigned)ceil_div(p.total, kBlock)), dim3(kBlock), 0, s>>>(
      p, (TO *)out);
  ESR_LAUNCHED();
  return ESR_OK;
}

template <typename TI>
static int dispatch_ew_out(int out_dtype, const EwParams &p, void *out, hipStream_t s) {
  switch (out_dtype) {
  case ESR_F32: return launch_ew<TI, float>(p, out, s);
  case ESR_F64: return launch_ew<TI, double>(p, out, s);
  case ESR_I32: return launch_ew<TI, int32_t>(p, out, s);
  case ESR_I64: return launch_ew<TI, int64_t>(p, out, s);
  case ESR_U8: return launch_ew<TI, uint8_t>(p, out, s);
  default:
    set_error("esr_ewise: unsupported out dtype %d", out_dtype);
    return ESR_ERR_UNSUPPORTED;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(T *__restrict__ dst, int64_t n, T v) {
  const int64_t tile = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t base = tile * (4 * kBlock) + threadIdx.x;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t i = base + k * kBlock;
    if (i < n) dst[i] = v;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_gemv_rows(const T *__restrict__ A, int64_t n,
                                                     int64_t kdim, int64_t lda, int64_t ldk,
                                                     const T *__restrict__ x, int64_t incx,
                                                     T *__restrict__ y) {
  const int64_t tile = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t r = tile * kBlock + threadIdx.x;
  if (r >= n) return;
  T acc = T(0);
  for (int64_t k = 0; k < kdim; ++k) acc = acc + A[r * lda + k * ldk] * x[k * incx];
  y[r] = acc;
}

} // namespace esr

using namespace esr;

// ============================================================================
// ABI
// ============================================================================
ESR_API int esr_gather(int dtype, const void *src, int64_t n_src, int64_t width,
                       const int32_t *idx, int64_t n_out, void *dst, void *stream) {
  ESR_GUARD_BEGIN
  const size_t es = dtype_size(dtype);
  ESR_REQUIRE(es > 0, "esr_gather: unknown dtype %d", dtype);
  ESR_REQUIRE(n_src >= 0 && n_out >= 0 && width >= 1, "esr_gather: bad sizes");
  if (n_out == 0) return ESR_OK;
  ESR_REQUIRE(src && idx && dst, "esr_gather: null pointer");
  ESR_REQUIRE(n_src > 0, "esr_gather: empty source with non-empty index");
  const int64_t row_bytes = (int64_t)es * width;
  const uintptr_t al = (uintptr_t)src | (uintptr_t)dst;
  hipStream_t s = as_stream(stream);
  if (row_bytes % 16 == 0 && al % 16 == 0)
    return launch_gather<uint4>(src, idx, dst, n_out, row_bytes / 16, s);
  if (row_bytes % 8 == 0 && al % 8 == 0)
    return launch_gather<uint2>(src, idx, dst, n_out, row_bytes / 8, s);
  if (row_bytes % 4 == 0 && al % 4 == 0)
    return launch_gather<uint32_t>(src, idx, dst, n_out, row_bytes / 4, s);
  if (row_bytes % 2 == 0 && al % 2 == 0)
    return launch_gather<uint16_t>(src, idx, dst, n_out, row_bytes / 2, s);
  return launch_gather<uint8_t>(src, idx, dst, n_out, row_bytes, s);
  ESR_GUARD_END
}

ESR_API int esr_csr_reduce(int op, int dtype, const void *vals, int64_t width,
                           const int32_t *rowptr, const int32_t *perm, int64_t n_rows,
                           const int32_t *long_rows, int64_t n_long, int keep_empty,
                           void *out, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(width >= 1 && n_rows >= 0 && n_long >= 0, "esr_csr_reduce: bad sizes");
  if (n_rows == 0) return ESR_OK;
  ESR_REQUIRE(rowptr && out, "esr_csr_reduce: null pointer");
  ESR_REQUIRE(n_long == 0 || long_rows, "esr_csr_reduce: null long_rows");
  hipStream_t s = as_stream(stream);
  switch (dtype) {
  case ESR_F32:
    return dispatch_csr_op<float>(op, vals, width, rowptr, perm, n_rows, long_rows, n_long,
                                  keep_empty, out, s);
  case ESR_F64:
    return dispatch_csr_op<double>(op, vals, width, rowptr, perm, n_rows, long_rows, n_long,
                                   keep_empty, out, s);
  case ESR_I32:
    return dispatch_csr_op<int32_t>(op, vals, width, rowptr, perm, n_rows, long_rows, n_long,
                                    keep_empty, out, s);
  case ESR_I64:
    return dispatch_csr_op<int64_t>(op, vals, width, rowptr, perm, n_rows, long_rows, n_long,
                                    keep_empty, out, s);
  default:
    set_error("esr_csr_reduce: unsupported dtype %d", dtype);
    return ESR_ERR_UNSUPPORTED;
  }
  ESR_GUARD_END
}

ESR_API int esr_aggregate_ws_bytes(int dtype, int64_t n, int64_t width, int64_t *bytes) {
  ESR_REQUIRE(bytes != nullptr && n >= 0 && width >= 1, "esr_aggregate_ws_bytes: bad args");
  ESR_REQUIRE(dtype_size(dtype) > 0, "esr_aggregate_ws_bytes: unknown dtype");
  *bytes = agg_blocks(n) * width * (int64_t)dtype_size(dtype);
  return ESR_OK;
}

template <typename T>
static int run_agg(int op, const void *x, int64_t n, int64_t width, double p, void *out,
                   void *ws, hipStream_t s) {
  const int64_t nb = agg_blocks(n);
  k_agg_stage1<T><<<dim3((unsigned)nb, (unsigned)width), dim3(kBlock), 0, s>>>(
      (const T *)x, n, width, op, p, (T *)ws);
  ESR_LAUNCHED();
  k_agg_stage2<T><<<dim3((unsigned)width), dim3(kBlock), 0, s>>>((const T *)ws, nb, op, p,
                                                                 (T *)out);
  ESR_LAUNCHED();
  return ESR_OK;
}

ESR_API int esr_aggregate(int op, int dtype, const void *x, int64_t n, int64_t width, double p,
                          void *out, void *ws, int64_t ws_bytes, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(op >= ESR_AGG_SUM && op <= ESR_AGG_MIN, "esr_aggregate: unknown op %d", op);
  ESR_REQUIRE(n >= 0 && width >= 1, "esr_aggregate: bad sizes");
  ESR_REQUIRE(out && ws && (x || n == 0), "esr_aggregate: null pointer");
  int64_t need = 0;
  esr_aggregate_ws_bytes(dtype, n, width, &need);
  ESR_REQUIRE(ws_bytes >= need, "esr_aggregate: workspace %lld < %lld bytes",
              (long long)ws_bytes, (long long)need);
  hipStream_t s = as_stream(stream);
  switch (dtype) {
  case ESR_F32: return run_agg<float>(op, x, n, width, p, out, ws, s);
  case ESR_F64: return run_agg<double>(op, x, n, width, p, out, ws, s);
  case ESR_I32:
    ESR_REQUIRE(op != ESR_AGG_NORM, "esr_aggregate: norm needs floating point");
    return run_agg<int32_t>(op, x, n, width, p, out, ws, s);
  case ESR_I64:
    ESR_REQUIRE(op != ESR_AGG_NORM, "esr_aggregate: norm needs floating point");
    return run_agg<int64_t>(op, x, n, width, p, out, ws, s);
  default:
    set_error("esr_aggregate: unsupported dtype %d", dtype);
    return ESR_ERR_UNSUPPORTED;
  }
  ESR_GUARD_END
}

// ----------------------------------------------------------------------------
// esr_aggregate_noalloc: the same two stages with the stage-1 partials in a
// device array the library holds (one copy per GPU, for the life of the
// process), for callers that cannot allocate.
// ----------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[ESR_AGG_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (each GPU holds its own copy of the array at its own address; the
// array's name in host code is only a stand-in for the lookup).
int scratch_on_current_device(void **out) {
  int dev = -1;
  ESR_HIP(hipGetDevice(&dev));
  ESR_REQUIRE(dev >= 0 && dev < kMaxDevices, "esr_aggregate_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[dev] == nullptr) {
    void *p = nullptr;
    ESR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ESR_REQUIRE(p != nullptr, "esr_aggregate_noalloc: no device address for the scratch array");
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return ESR_OK;
}
} // namespace

ESR_API int esr_aggregate_noalloc(int op, int dtype, const void *x, int64_t n, int64_t width,
                                  double p, void *out, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 0 && width >= 1, "esr_aggregate_noalloc: bad sizes");
  ESR_REQUIRE(dtype_size(dtype) > 0, "esr_aggregate_noalloc: unknown dtype %d", dtype);
  int64_t need = 0;
  esr_aggregate_ws_bytes(dtype, n, width, &need);
  ESR_REQUIRE(need <= (int64_t)ESR_AGG_SCRATCH_BYTES,
              "esr_aggregate_noalloc: %lld x %lld values need %lld scratch bytes, more than the "
              "%d the library holds (use esr_aggregate with a workspace)",
              (long long)n, (long long)width, (long long)need, (int)ESR_AGG_SCRATCH_BYTES);
  void *ws = nullptr;
  const int rc = scratch_on_current_device(&ws);
  if (rc != ESR_OK) return rc;
  return esr_aggregate(op, dtype, x, n, width, p, out, ws, need, stream);
  ESR_GUARD_END
}

namespace esr {
namespace {
// Dense 1-d copy / add / sub / mul / div in one floating dtype: one element
// per thread, inputs with stride 1, stride 0 (a one-element tensor) or a
// scalar (null pointer). The streaming case of the general kernel above
// without its per-element shape walk and operator switch.
template <typename T, int OP>
__global__ __launch_bounds__(kBlock) void k_dense(int64_t n, const T *__restrict__ a, int64_t sa,
                                                 T ca, const T *__restrict__ b, int64_t sb, T cb,
                                                 T *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const T x = a ? a[i * sa] : ca;
  if constexpr (OP == ESR_EW_COPY) {
    out[i] = x;
  } else {
    const T y = b ? b[i * sb] : cb;
    if constexpr (OP == ESR_EW_ADD) out[i] = x + y;
    else if constexpr (OP == ESR_EW_SUB) out[i] = x - y;
    else if constexpr (OP == ESR_EW_MUL) out[i] = x * y;
    else out[i] = x / y;
  }
}

template <typename T>
int launch_dense(const EwParams &p, void *out, hipStream_t s) {
  const unsigned g = (unsigned)ceil_div(p.total, kBlock);
  const T *a = (const T *)p.in[0], *b = (const T *)p.in[1];
  const int64_t sa = p.istr[0][0], sb = p.istr[1][0];
  const T ca = (T)p.sc[0], cb = (T)p.sc[1];
  switch (p.op) {
  case ESR_EW_COPY: k_dense<T, ESR_EW_COPY><<<g, kBlock, 0, s>>>(p.total, a, sa, ca, b, sb, cb, (T *)out); break;
  case ESR_EW_ADD: k_dense<T, ESR_EW_ADD><<<g, kBlock, 0, s>>>(p.total, a, sa, ca, b, sb, cb, (T *)out); break;
  case ESR_EW_SUB: k_dense<T, ESR_EW_SUB><<<g, kBlock, 0, s>>>(p.total, a, sa, ca, b, sb, cb, (T *)out); break;
  case ESR_EW_MUL: k_dense<T, ESR_EW_MUL><<<g, kBlock, 0, s>>>(p.total, a, sa, ca, b, sb, cb, (T *)out); break;
  default: k_dense<T, ESR_EW_DIV><<<g, kBlock, 0, s>>>(p.total, a, sa, ca, b, sb, cb, (T *)out); break;
  }
  ESR_LAUNCHED();
  return ESR_OK;
}
} // namespace
} // namespace esr

ESR_API int esr_ewise(int op, int ndim, const int64_t *shape, int out_dtype, void *out,
                      const int64_t *out_strides, int in_dtype, int n_in,
                      void *const *in_ptrs, const int64_t *in_strides, const double *scalars,
                      void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(ndim >= 0 && ndim <= ESR_EW_MAXDIM, "esr_ewise: ndim %d out of range", ndim);
  ESR_REQUIRE(n_in >= 1 && n_in <= 3, "esr_ewise: n_in %d out of range", n_in);
  const int want = op >= ESR_EW_WHERE ? 3 : (op >= ESR_EW_ADD ? 2 : 1);
  ESR_REQUIRE(n_in == want, "esr_ewise: op %d takes %d inputs, got %d", op, want, n_in);
  EwParams p{};
  p.ndim = ndim == 0 ? 1 : ndim;
  p.op = op;
  p.n_in = n_in;
  p.total = 1;
  for (int d = 0; d < ESR_EW_MAXDIM; ++d) {
    p.shape[d] = 1;
    p.ostr[d] = 0;
    for (int k = 0; k < 3; ++k) p.istr[k][d] = 0;
  }
  for (int d = 0; d < ndim; ++d) {
    ESR_REQUIRE(shape[d] >= 0, "esr_ewise: negative extent");
    p.shape[d] = shape[d];
    p.ostr[d] = out_strides[d];
    p.total *= shape[d];
  }
  for (int k = 0; k < n_in; ++k) {
    p.in[k] = in_ptrs[k];
    p.sc[k] = scalars ? scalars[k] : 0.0;
    if (in_ptrs[k])
      for (int d = 0; d < ndim; ++d) p.istr[k][d] = in_strides[k * ndim + d];
  }
  if (ndim == 0) {
    p.ostr[0] = 1;
    for (int k = 0; k < 3; ++k) p.istr[k][0] = 0;
  }
  if (p.total == 0) return ESR_OK;
  ESR_REQUIRE(out != nullptr, "esr_ewise: null out");
  // dense 1-d fast path: out stride 1, every tensor input stride 1 or 0
  p.contiguous = (p.ndim == 1 && p.ostr[0] == 1);
  for (int k = 0; k < n_in && p.contiguous; ++k)
    if (p.in[k] && p.istr[k][0] != 1 && p.istr[k][0] != 0) p.contiguous = 0;
  hipStream_t s = as_stream(stream);
  const bool dense_op = op == ESR_EW_COPY || op == ESR_EW_ADD || op == ESR_EW_SUB ||
                        op == ESR_EW_MUL || op == ESR_EW_DIV;
  if (p.contiguous && dense_op && in_dtype == out_dtype && p.total < (1ll << 40)) {
    if (in_dtype == ESR_F64) return launch_dense<double>(p, out, s);
    if (in_dtype == ESR_F32) return launch_dense<float>(p, out, s);
  }
  if (op == ESR_EW_WHERE) {
    // condition is u8; value inputs in_dtype
    switch (in_dtype) {
    case ESR_F32: return launch_ew<float, float>(p, out, s);
    case ESR_F64: return launch_ew<double, double>(p, out, s);
    case ESR_I32: return launch_ew<int32_t, int32_t>(p, out, s);
    case ESR_I64: return launch_ew<int64_t, int64_t>(p, out, s);
    case ESR_U8: return launch_ew<uint8_t, uint8_t>(p, out, s);
    default: break;
    }
  } else {
    switch (in_dtype) {
    case ESR_F32: return dispatch_ew_out<float>(out_dtype, p, out, s);
    case ESR_F64: return dispatch_ew_out<double>(out_dtype, p, out, s);
    case ESR_I32: return dispatch_ew_out<int32_t>(out_dtype, p, out, s);
    case ESR_I64: return dispatch_ew_out<int64_t>(out_dtype, p, out, s);
    case ESR_U8: return dispatch_ew_out<uint8_t>(out_dtype, p, out, s);
    default: break;
    }
  }
  set_error("esr_ewise: unsupported input dtype %d", in_dtype);
  return ESR_ERR_UNSUPPORTED;
  ESR_GUARD_END
}

ESR_API int esr_fill(int dtype, void *dst, int64_t n, double value, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 0, "esr_fill: negative size");
  if (n == 0) return ESR_OK;
  ESR_REQUIRE(dst != nullptr, "esr_fill: null pointer");
  hipStream_t s = as_stream(stream);
  const dim3 grid((unsigned)ceil_div(n, 4 * kBlock));
  switch (dtype) {
  case ESR_F32: k_fill<float><<<grid, kBlock, 0, s>>>((float *)dst, n, (float)value); break;
  case ESR_F64: k_fill<double><<<grid, kBlock, 0, s>>>((double *)dst, n, value); break;
  case ESR_I32: k_fill<int32_t><<<grid, kBlock, 0, s>>>((int32_t *)dst, n, (int32_t)value); break;
  case ESR_I64: k_fill<int64_t><<<grid, kBlock, 0, s>>>((int64_t *)dst, n, (int64_t)value); break;
  case ESR_U8: k_fill<uint8_t><<<grid, kBlock, 0, s>>>((uint8_t *)dst, n, (uint8_t)(value != 0.0)); break;
  default: set_error("esr_fill: unsupported dtype %d", dtype); return ESR_ERR_UNSUPPORTED;
  }
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

ESR_API int esr_gemv_rows(int dtype, const void *A, int64_t n, int64_t kdim, int64_t lda,
                          int64_t ldk, const void *x, int64_t incx, void *y, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 0 && kdim >= 0, "esr_gemv_rows: bad sizes");
  if (n == 0) return ESR_OK;
  ESR_REQUIRE(y && (kdim == 0 || (A && x)), "esr_gemv_rows: null pointer");
  hipStream_t s = as_stream(stream);
  const dim3 grid((unsigned)ceil_div(n, kBlock));
  switch (dtype) {
  case ESR_F32:
    k_gemv_rows<float><<<grid, kBlock, 0, s>>>((const float *)A, n, kdim, lda, ldk,
                                               (const float *)x, incx, (float *)y);
    break;
  case ESR_F64:
    k_gemv_rows<double><<<grid, kBlock, 0, s>>>((const double *)A, n, kdim, lda, ldk,
                                                (const double *)x, incx, (double *)y);
    break;
  default: set_error("esr_gemv_rows: unsupported dtype %d", dtype); return ESR_ERR_UNSUPPORTED;
  }
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}
