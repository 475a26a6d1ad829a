// Device helpers shared by the Krylov-solver kernels: grid-stride row loops,
// the Linsys row product, and fence-free deterministic grid reductions.
//
// Reductions. The MI355X's L2 caches are per XCD and not coherent with each
// other, so an agent-scope release fence -- what a "the last workgroup to
// arrive finishes the sum" reduction needs before its arrival atomic --
// writes back and invalidates the XCD's whole L2 (buffer_wbl2 / buffer_inv),
// once per workgroup. Here a reduction kernel instead stores one partial per
// workgroup with plain stores, and the next kernel on the stream (ordered
// after every store of the previous one) sums all partials in every
// workgroup, in one fixed order, so every workgroup holds the same bits. The
// grid of such kernels is capped at kMaxGrid workgroups (rows are visited
// with a grid-stride loop) so that re-summing the partials stays cheap.
#pragma once
#include "esr_common.h"

namespace esr {

constexpr int kRB = 256;        // threads per workgroup of the solver kernels
constexpr int kMaxGrid = 1024;  // workgroups of a grid-stride solver kernel

inline unsigned red_grid(int64_t n) {
  const int64_t g = ceil_div(n, kRB);
  return (unsigned)(g < kMaxGrid ? (g > 0 ? g : 1) : kMaxGrid);
}

// Rows of a grid-stride kernel: i = blockIdx.x * kRB + threadIdx.x + k * stride.
// (A contiguous chunk of rows per workgroup, which keeps a row's neighbours
// in its XCD's L2, measured slower: Jacobi-PCG at 4.2 M vertices 1.62 vs
// 1.54 s, the 8 XCDs then stream from 8 distant regions.)
#define ESR_GRID_ROWS(i, n)                                                    \
  for (int64_t i = (int64_t)blockIdx.x * kRB + threadIdx.x; i < (n);           \
       i += (int64_t)gridDim.x * kRB)

__device__ __forceinline__ double kr_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Workgroup sums of NV per-thread values, written to part[k * G + blockIdx.x]
// (G = gridDim.x) by thread 0. Fixed combination order.
template <int NV>
__device__ __forceinline__ void store_partials(const double (&v)[NV], double *part) {
  __shared__ double sh[NV][kRB / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double s = kr_wave_sum(v[k]);
    if (lane == 0) sh[k][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int w = 0; w < kRB / 64; ++w) s += sh[k][w];
      part[(int64_t)k * gridDim.x + blockIdx.x] = s;
    }
  }
  __syncthreads();  // sh may be reused by a following reduction
}

// out[k] = sum_{i < cnt} part[k * stride + first + i] for every thread of the
// workgroup, in a fixed order (identical in every workgroup).
template <int NV>
__device__ __forceinline__ void sum_partials(const double *part, int64_t stride, int64_t first,
                                             int64_t cnt, double (&out)[NV]) {
  __shared__ double sh[NV][kRB / 64];
  __shared__ double res[NV];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < cnt; i += kRB) s += part[k * stride + first + i];
    s = kr_wave_sum(s);
    if (lane == 0) sh[k][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double t = 0.0;
      for (int w = 0; w < kRB / 64; ++w) t += sh[k][w];
      res[k] = t;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) out[k] = res[k];
  __syncthreads();  // res / sh may be reused by the caller's next reduction
}

// Linsys row (easier/numeric/linsys.py:23-24): the off-diagonal CSR sum of
// row r folded from 0 in entry order, plus the diagonal term. The first 8
// entries are evaluated as independent slots (clamped positions, so every
// load is valid and all of them are in flight together) and folded in only
// when they exist: the same left-to-right order as the plain loop.
template <typename T, typename Src>
__device__ __forceinline__ T linsys_row(int64_t r, const int32_t *__restrict__ rowptr,
                                        const int32_t *__restrict__ col,
                                        const T *__restrict__ Ae, const T *__restrict__ Av,
                                        Src src) {
  T s = T(0);
  const int32_t b = rowptr[r], e = rowptr[r + 1];
  if (e > b) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int32_t j = b + k;
      const int32_t jc = j < e ? j : e - 1;
      const T v = src(col[jc]) * Ae[jc];
      s = j < e ? s + v : s;
    }
    for (int32_t j = b + 8; j < e; ++j) s = s + src(col[j]) * Ae[j];
  }
  return s + src(r) * Av[r];
}

// The same row with the slot count chosen per wavefront: K = 4, 6 or 8
// independent slots, the smallest that covers the longest row of the
// lanes executing it (wave-uniform, from ballots), so rows of a mesh whose
// vertices have 6 neighbours do not issue the loads of two empty slots. The
// fold is the same left-to-right sum of the row's entries.
template <int K, typename T, typename Src>
__device__ __forceinline__ T linsys_slots(int32_t b, int32_t e, const int32_t *__restrict__ col,
                                          const T *__restrict__ Ae, Src src) {
  T s = T(0);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int32_t j = b + k;
    const int32_t jc = j < e ? j : (e > b ? e - 1 : b);
    const T v = src(col[jc]) * Ae[jc];
    s = j < e ? s + v : s;
  }
  for (int32_t j = b + K; j < e; ++j) s = s + src(col[j]) * Ae[j];
  return s;
}

// ---------------------------------------------------------------------------
// Sliced ELLPACK with slices of one wavefront (SELL-64; Kreutzer, Hager,
// Wellein, Fehske and Bishop, SIAM J. Sci. Comput. 36(5), 2014, with sigma =
// 1: rows keep their order). Rows 64 s .. 64 s + 63 form slice s; its width
// w_s is the longest of those rows; entry k of row r = 64 s + l sits at
// soff[s] + 64 k + l (soff[s + 1] - soff[s] = 64 w_s), so each slot is one
// fully used, contiguous load per wavefront where a CSR row spreads a
// wavefront's loads over ~6 rows' worth of cache lines per instruction.
// cnt[r] (uint8) is the row's length; padded entries hold value 0 and a
// column inside the matrix, and are not folded. Column indices are int16
// offsets from the row (when every one fits, e.g. a mesh renumbered for
// locality) or int32 absolute indices. The row is folded exactly like
// linsys_row: from 0, entries in CSR order, diagonal term last.
__device__ __forceinline__ int64_t sell_col(int64_t r, const int16_t *__restrict__ c,
                                            int64_t e) {
  return r + (int64_t)c[e];
}
__device__ __forceinline__ int64_t sell_col(int64_t, const int32_t *__restrict__ c, int64_t e) {
  return (int64_t)c[e];
}

// K independent slots (K >= 1), slot positions clamped to the slice width
// w (every load inside the slice), folded only where k < c.
template <int K, typename T, typename C, typename Src>
__device__ __forceinline__ T sell_slots(int64_t r, int64_t e0, int w, int c,
                                        const C *__restrict__ colv, const T *__restrict__ val,
                                        Src src) {
  T s = T(0);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int kk = k < w ? k : w - 1;
    const int64_t e = e0 + 64 * (int64_t)kk;
    const T v = src(sell_col(r, colv, e)) * val[e];
    s = k < c ? s + v : s;
  }
  for (int k = K; k < w; ++k) {
    const int64_t e = e0 + 64 * (int64_t)k;
    const T v = src(sell_col(r, colv, e)) * val[e];
    s = k < c ? s + v : s;
  }
  return s;
}

// Row r of the Linsys from its SELL-64 form. Every active lane of the
// wavefront must hold a row of the same slice (rows visited 64 at a time,
// wavefront-aligned), so the slice's offset and width are wave-uniform.
template <typename T, typename C, typename Src>
__device__ __forceinline__ T sell_row(int64_t r, const int64_t *__restrict__ soff,
                                      const uint8_t *__restrict__ cnt,
                                      const C *__restrict__ colv, const T *__restrict__ val,
                                      const T *__restrict__ Av, Src src) {
  const int s = __builtin_amdgcn_readfirstlane((int)(r >> 6));
  const int64_t base = soff[s];
  const int w = (int)((soff[s + 1] - base) >> 6);
  const int64_t e0 = base + (r & 63);
  const int c = cnt[r];
  T y = T(0);
  if (w > 6)
    y = sell_slots<8, T>(r, e0, w, c, colv, val, src);
  else if (w > 4)
    y = sell_slots<6, T>(r, e0, w, c, colv, val, src);
  else if (w > 0)
    y = sell_slots<4, T>(r, e0, w, c, colv, val, src);
  return y + src(r) * Av[r];
}

// A Linsys operator as a kernel argument: row(r, src) folds row r exactly as
// linsys_row does. CsrRows reads the CSR; SellRows the SELL-64 form (rows
// must be visited 64 at a time, wavefront-aligned).
template <typename T> struct CsrRows {
  const int32_t *rowptr;
  const int32_t *col;
  const T *Ae;
  const T *Av;
  template <typename Src> __device__ __forceinline__ T row(int64_t r, Src src) const {
    return linsys_row<T>(r, rowptr, col, Ae, Av, src);
  }
};
template <typename T, typename C> struct SellRows {
  const int64_t *soff;
  const uint8_t *cnt;
  const C *colv;
  const T *val;
  const T *Av;
  template <typename Src> __device__ __forceinline__ T row(int64_t r, Src src) const {
    return sell_row<T>(r, soff, cnt, colv, val, Av, src);
  }
};

template <typename T, typename Src>
__device__ __forceinline__ T linsys_row_wave(int64_t r, const int32_t *__restrict__ rowptr,
                                             const int32_t *__restrict__ col,
                                             const T *__restrict__ Ae, const T *__restrict__ Av,
                                             Src src) {
  const int32_t b = rowptr[r], e = rowptr[r + 1];
  T s = T(0);
  if (__ballot(e - b > 6)) {
    if (e > b) s = linsys_slots<8, T>(b, e, col, Ae, src);
  } else if (__ballot(e - b > 4)) {
    if (e > b) s = linsys_slots<6, T>(b, e, col, Ae, src);
  } else {
    if (e > b) s = linsys_slots<4, T>(b, e, col, Ae, src);
  }
  return s + src(r) * Av[r];
}

} // namespace esr
