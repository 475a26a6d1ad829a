// Fused Krylov iterations over a CSR Linsys operator.
//
// The operator is Linsys (easier/numeric/linsys.py:23-24):
//   y[r] = (sum_{j in row r, in edge order} x[col[j]] * Ae[j]) + x[r] * Av[r]
// with the row sum folded from 0 left to right, which is what the reference's
// Reducer computes for the renumbered, row-sorted edge list.
//
// Dot products are accumulated in double with a fixed tree and finalised by
// the last workgroup to finish (an integer arrival counter, no float atomics),
// so every run produces the same bits. Solver scalars stay on the device.
#include "esr_common.h"
#include "esr_solver.h"

namespace esr {
namespace {

constexpr int kB = 256;          // threads per workgroup
constexpr int kRowsPerThread = 4;
constexpr int kTile = kB * kRowsPerThread;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Sums NV per-thread values over the workgroup, stores the NV workgroup sums
// at part[k * nb + b]; the last workgroup to arrive sums part[k * nb + 0..nb)
// in index order and returns true with the totals in out[] (thread 0).
template <int NV>
__device__ bool grid_sum_last(double (&v)[NV], double *__restrict__ part,
                              uint32_t *counter, int64_t nb, int64_t b,
                              double (&out)[NV]) {
  __shared__ double sh[NV][kB / 64];
  __shared__ int last;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double s = wave_sum(v[k]);
    if (lane == 0) sh[k][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int w = 0; w < kB / 64; ++w) s += sh[k][w];
      __hip_atomic_store(part + k * nb + b, s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __threadfence();
    const uint32_t prev = atomicAdd(counter, 1u);
    last = (prev == (uint32_t)(nb - 1));
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < nb; i += kB)
      s += __hip_atomic_load(part + k * nb + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s = wave_sum(s);
    __syncthreads();
    if (lane == 0) sh[k][wid] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int w = 0; w < kB / 64; ++w) t += sh[k][w];
      out[k] = t;
    }
  }
  if (threadIdx.x == 0) *counter = 0u;
  return true;
}

template <typename T, typename Src>
__device__ __forceinline__ T spmv_row(int64_t r, const int32_t *__restrict__ rowptr,
                                      const int32_t *__restrict__ col,
                                      const T *__restrict__ Ae, const T *__restrict__ Av,
                                      Src src) {
  T s = T(0);
  const int32_t b = rowptr[r], e = rowptr[r + 1];
  for (int32_t j = b; j < e; ++j) s = s + src(col[j]) * Ae[j];
  return s + src(r) * Av[r];
}

// ---------------------------------------------------------------------------
// PCG
// ---------------------------------------------------------------------------
// Init (cg.py:25-29): r = b - A x; z = M r; p = z; bnorm = ||b||; rz = r.z;
// tol = max(rtol*bnorm, atol) if atol else rtol*bnorm (cg.py:112-115).
template <typename T>
__global__ __launch_bounds__(kB) void k_pcg_init(
    int64_t n, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const T *__restrict__ Ae, const T *__restrict__ Av, const T *__restrict__ diag,
    const T *__restrict__ bvec, const T *__restrict__ x, T *__restrict__ r,
    T *__restrict__ z, T *__restrict__ p, PcgState *st, double *part) {
  const int64_t b = xcd_remap(blockIdx.x, gridDim.x);
  double acc[2] = {0.0, 0.0};
  for (int k = 0; k < kRowsPerThread; ++k) {
    const int64_t i = b * kTile + k * kB + threadIdx.x;
    if (i >= n) break;
    const T ax = spmv_row<T>(i, rowptr, col, Ae, Av, [&](int64_t c) { return x[c]; });
    const T ri = bvec[i] - ax;
    const T zi = diag ? ri / diag[i] : ri;
    r[i] = ri;
    z[i] = zi;
    p[i] = zi;
    acc[0] += (double)(ri * zi);
    acc[1] += (double)(bvec[i] * bvec[i]);
  }
  double tot[2];
  if (grid_sum_last<2>(acc, part, &st->counter[0], gridDim.x, b, tot) && threadIdx.x == 0) {
    st->rz_cur = tot[0];
    st->rz_next = tot[0];
    st->bnorm = cr_sqrt(tot[1]);
    double tol = st->rtol * st->bnorm;
    if (st->atol > 0.0 && st->atol > tol) tol = st->atol;
    st->tol = tol;
    st->k_cur = 0;
    st->k_next = 0;
    st->done = 0;
    st->iters = 0;
  }
}

// Step part 1 (cg.py:44-47): Ap = A p; pAp = p.Ap; alpha = rz / pAp.
template <typename T>
__global__ __launch_bounds__(kB) void k_pcg_spmv(
    int64_t n, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const T *__restrict__ Ae, const T *__restrict__ Av, const T *__restrict__ p,
    T *__restrict__ Ap, PcgState *st, double *part) {
  if (st->done) return;
  const int64_t b = xcd_remap(blockIdx.x, gridDim.x);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->k_cur = st->k_next;   // nobody reads these in this kernel
    st->rz_cur = st->rz_next;
  }
  double acc[1] = {0.0};
  for (int k = 0; k < kRowsPerThread; ++k) {
    const int64_t i = b * kTile + k * kB + threadIdx.x;
    if (i >= n) break;
    const T y = spmv_row<T>(i, rowptr, col, Ae, Av, [&](int64_t c) { return p[c]; });
    Ap[i] = y;
    acc[0] += (double)(p[i] * y);
  }
  double tot[1];
  if (grid_sum_last<1>(acc, part, &st->counter[1], gridDim.x, b, tot) && threadIdx.x == 0) {
    st->pAp = tot[0];
    st->alpha = st->rz_next / tot[0];
  }
}

// Step part 2 (cg.py:48-52) fused with the first half of Update (cg.py:67):
// x += alpha p; r -= alpha Ap; z = M r; rr = r.r; rz = r.z.
template <typename T>
__global__ __launch_bounds__(kB) void k_pcg_xr(int64_t n, const T *__restrict__ diag,
                                              T *__restrict__ x, T *__restrict__ r,
                                              T *__restrict__ z, const T *__restrict__ p,
                                              const T *__restrict__ Ap, PcgState *st,
                                              double *part) {
  if (st->done) return;
  const int64_t b = xcd_remap(blockIdx.x, gridDim.x);
  const T alpha = (T)st->alpha;
  double acc[2] = {0.0, 0.0};
  for (int k = 0; k < kRowsPerThread; ++k) {
    const int64_t i = b * kTile + k * kB + threadIdx.x;
    if (i >= n) break;
    const T pi = p[i];
    const T ri = r[i] - alpha * Ap[i];
    x[i] = x[i] + alpha * pi;
    r[i] = ri;
    const T zi = diag ? ri / diag[i] : ri;
    z[i] = zi;
    acc[0] += (double)(ri * ri);
    acc[1] += (double)(ri * zi);
  }
  double tot[2];
  if (grid_sum_last<2>(acc, part, &st->counter[2], gridDim.x, b, tot) && threadIdx.x == 0) {
    st->rr = tot[0];
    st->rz_new = tot[1];
    st->rnorm = cr_sqrt(tot[0]);
  }
}

// Loop test (cg.py:126-129) and Update (cg.py:67-71): p = z + beta p.
template <typename T>
__global__ __launch_bounds__(kB) void k_pcg_p(int64_t n, const T *__restrict__ z,
                                             T *__restrict__ p, PcgState *st) {
  if (st->done) return;
  const double rnorm = st->rnorm;
  const int64_t k = st->k_cur;
  const bool converged = (rnorm == rnorm) && rnorm <= st->tol;
  const bool capped = st->maxiter >= 0 && k >= st->maxiter;
  const bool nan_stop = rnorm != rnorm && st->maxiter < 0;
  if (converged || capped || nan_stop) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->done = 1;
      st->iters = k;
    }
    return;
  }
  const T beta = (T)(st->rz_new / st->rz_cur);
  const int64_t b = xcd_remap(blockIdx.x, gridDim.x);
  for (int kk = 0; kk < kRowsPerThread; ++kk) {
    const int64_t i = b * kTile + kk * kB + threadIdx.x;
    if (i >= n) break;
    p[i] = z[i] + beta * p[i];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->k_next = k + 1;
    st->rz_next = st->rz_new;
    st->beta = (double)beta;
  }
}

// ---------------------------------------------------------------------------
// GMRES Arnoldi cycle
// ---------------------------------------------------------------------------
// ||b|| (Init, gmres.py:36-37)
template <typename T>
__global__ __launch_bounds__(kB) void k_gm_bnorm(int64_t n, const T *__restrict__ bvec,
                                                GmresState *st, double *part) {
  const int64_t b = xcd_remap(blockIdx.x, gridDim.x);
  double acc[1] = {0.0};
  for (int k = 0; k < kRowsPerThread; ++k) {
    const int64_t i = b * kTile + k * kB + threadIdx.x;
    if (i >= n) break;
    acc[0] += (double)(bvec[i] * bvec[i]);
  }
  double tot[1];
  if (grid_sum_last<1>(acc, part, &st->counter[0], gridDim.x, b, tot) && threadIdx.x == 0)
    st->bnorm = cr_sqrt(tot[0]);
}

// UpdateRnorm (gmres.py:22-25): r = b - A x; rnorm = ||r||. Clears the
// cycle's breakdown flag.
template <typename T>
__global__ __launch_bounds__(kB) void k_gm_residual(
    int64_t n, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const T *__restrict__ Ae, const T *__restrict__ Av, const T *__restrict__ bvec,
    const T *__restrict__ x, T *__restrict__ r, GmresState *st, double *part) {
  const int64_t b = xcd_remap(blockIdx.x, gridDim.x);
  double acc[1] = {0.0};
  for (int k = 0; k < kRowsPerThread; ++k) {
    const int64_t i = b * kTile + k * kB + threadIdx.x;
    if (i >= n) break;
    const T ri = bvec[i] - spmv_row<T>(i, rowptr, col, Ae, Av, [&](int64_t c) { return x[c]; });
    r[i] = ri;
    acc[0] += (double)(ri * ri);
  }
  double tot[1];
  if (grid_sum_last<1>(acc, part, &st->counter[1], gridDim.x, b, tot) && threadIdx.x == 0) {
    st->rnorm = cr_sqrt(tot[0]);
    st->broken = 0;
    st->jend = -1;
  }
}

// InitV/UpdateV + InitW (gmres.py:48, 64, 120): V_j = src / s (s = rnorm for
// j = 0, H[j, j-1] otherwise) and w = A(M(V_j)), M = identity or r / diag.
template <typename T>
__global__ __launch_bounds__(kB) void k_gm_spmv(
    int64_t n, int j, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const T *__restrict__ Ae, const T *__restrict__ Av, const T *__restrict__ diag,
    const T *__restrict__ src, T *__restrict__ Vj, T *__restrict__ w, GmresState *st) {
  if (st->broken) return;
  const T s = (T)(j == 0 ? st->rnorm : st->H[j * kGmresMaxRestart + (j - 1)]);
  const int64_t b = xcd_remap(blockIdx.x, gridDim.x);
  auto mv = [&](int64_t c) {
    const T v = src[c] / s;
    return diag ? v / diag[c] : v;
  };
  for (int k = 0; k < kRowsPerThread; ++k) {
    const int64_t i = b * kTile + k * kB + threadIdx.x;
    if (i >= n) break;
    Vj[i] = src[i] / s;
    w[i] = spmv_row<T>(i, rowptr, col, Ae, Av, mv);
  }
}

// Modified Gram-Schmidt step i of Arnoldi step j (gmres.py:79, 105):
// apply the pending w -= H[i-1, j] V_{i-1}, then H[i, j] = V_i . w.
template <typename T>
__global__ __launch_bounds__(kB) void k_gm_mgs(int64_t n, int i, int j,
                                              const T *__restrict__ Vprev,
                                              const T *__restrict__ Vi, T *__restrict__ w,
                                              GmresState *st, double *part) {
  if (st->broken) return;
  const int64_t b = xcd_remap(blockIdx.x, gridDim.x);
  const T hprev = i > 0 ? (T)st->H[(i - 1) * kGmresMaxRestart + j] : T(0);
  double acc[1] = {0.0};
  for (int k = 0; k < kRowsPerThread; ++k) {
    const int64_t r = b * kTile + k * kB + threadIdx.x;
    if (r >= n) break;
    T wr = w[r];
    if (i > 0) {
      wr = wr - hprev * Vprev[r];
      w[r] = wr;
    }
    acc[0] += (double)(Vi[r] * wr);
  }
  double tot[1];
  if (grid_sum_last<1>(acc, part, &st->counter[2], gridDim.x, b, tot) && threadIdx.x == 0)
    st->H[i * kGmresMaxRestart + j] = (double)(T)tot[0];
}

// Last MGS update and NormW (gmres.py:90, 210-213): w -= H[j, j] V_j;
// H[j+1, j] = ||w||; breakdown when it is below 1e-15.
template <typename T>
__global__ __launch_bounds__(kB) void k_gm_norm(int64_t n, int j, const T *__restrict__ Vj,
                                               T *__restrict__ w, GmresState *st,
                                               double *part) {
  if (st->broken) return;
  const int64_t b = xcd_remap(blockIdx.x, gridDim.x);
  const T h = (T)st->H[j * kGmresMaxRestart + j];
  double acc[1] = {0.0};
  for (int k = 0; k < kRowsPerThread; ++k) {
    const int64_t r = b * kTile + k * kB + threadIdx.x;
    if (r >= n) break;
    const T wr = w[r] - h * Vj[r];
    w[r] = wr;
    acc[0] += (double)(wr * wr);
  }
  double tot[1];
  if (grid_sum_last<1>(acc, part, &st->counter[3], gridDim.x, b, tot) && threadIdx.x == 0) {
    const double hn = (double)(T)cr_sqrt(tot[0]);
    st->H[(j + 1) * kGmresMaxRestart + j] = hn;
    st->jend = j;
    if (hn < 1e-15) st->broken = 1;
  }
}

// UpdateX (gmres.py:133-136): x += M(sum_{i<jc} V_i y_i).
template <typename T>
__global__ __launch_bounds__(kB) void k_gm_update_x(int64_t n, int jc, const T *__restrict__ V,
                                                   const T *__restrict__ y,
                                                   const T *__restrict__ diag,
                                                   T *__restrict__ x) {
  const int64_t b = xcd_remap(blockIdx.x, gridDim.x);
  for (int k = 0; k < kRowsPerThread; ++k) {
    const int64_t r = b * kTile + k * kB + threadIdx.x;
    if (r >= n) break;
    T dx = T(0);
    for (int i = 0; i < jc; ++i) dx = dx + V[(int64_t)i * n + r] * y[i];
    x[r] = x[r] + (diag ? dx / diag[r] : dx);
  }
}

// ---------------------------------------------------------------------------
// Mixed-precision iterative refinement (FP64 outer loop around FP32 CG)
// ---------------------------------------------------------------------------
struct IrState {
  double rnorm;
  double bnorm;
  uint32_t counter[2];
};

// r = b - A x in float64; writes r rounded to float32 and ||r|| (float64);
// with first != 0 also ||b||.
__global__ __launch_bounds__(kB) void k_ir_residual(
    int64_t n, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const double *__restrict__ Ae, const double *__restrict__ Av,
    const double *__restrict__ bvec, const double *__restrict__ x, float *__restrict__ r32,
    IrState *st, double *part, int first) {
  const int64_t b = xcd_remap(blockIdx.x, gridDim.x);
  double acc[2] = {0.0, 0.0};
  for (int k = 0; k < kRowsPerThread; ++k) {
    const int64_t i = b * kTile + k * kB + threadIdx.x;
    if (i >= n) break;
    const double ri =
        bvec[i] - spmv_row<double>(i, rowptr, col, Ae, Av, [&](int64_t c) { return x[c]; });
    r32[i] = (float)ri;
    acc[0] += ri * ri;
    acc[1] += bvec[i] * bvec[i];
  }
  double tot[2];
  if (grid_sum_last<2>(acc, part, &st->counter[0], gridDim.x, b, tot) && threadIdx.x == 0) {
    st->rnorm = cr_sqrt(tot[0]);
    if (first) st->bnorm = cr_sqrt(tot[1]);
  }
}

// x += (double) d; d = 0 (the next inner solve starts from zero).
__global__ __launch_bounds__(kB) void k_ir_update(int64_t n, double *__restrict__ x,
                                                 float *__restrict__ d) {
  const int64_t b = xcd_remap(blockIdx.x, gridDim.x);
  for (int k = 0; k < kRowsPerThread; ++k) {
    const int64_t i = b * kTile + k * kB + threadIdx.x;
    if (i >= n) break;
    x[i] = x[i] + (double)d[i];
    d[i] = 0.0f;
  }
}

inline unsigned grid_of(int64_t n) { return (unsigned)ceil_div(n, kTile); }

} // namespace
} // namespace esr

using namespace esr;

#define ESR_FLOAT_DISPATCH(dtype, ...)                                          \
  do {                                                                          \
    if ((dtype) == ESR_F64) {                                                   \
      using T = double;                                                         \
      __VA_ARGS__;                                                              \
    } else if ((dtype) == ESR_F32) {                                            \
      using T = float;                                                          \
      __VA_ARGS__;                                                              \
    } else {                                                                    \
      set_error("solver kernels need float32 or float64, got dtype %d", dtype); \
      return ESR_ERR_UNSUPPORTED;                                               \
    }                                                                           \
  } while (0)

ESR_API int esr_pcg_workspace(int64_t n, int64_t *state_bytes, int64_t *ws_bytes) {
  ESR_REQUIRE(n >= 1 && state_bytes && ws_bytes, "esr_pcg_workspace: bad args");
  *state_bytes = (int64_t)sizeof(PcgState);
  *ws_bytes = 2 * (int64_t)grid_of(n) * (int64_t)sizeof(double);
  return ESR_OK;
}

ESR_API int esr_pcg_init(int dtype, int64_t n, const int32_t *rowptr, const int32_t *col,
                         const void *Ae, const void *Av, const void *diag, const void *b,
                         void *x, void *r, void *z, void *p, void *state, void *ws,
                         double rtol, double atol, int64_t maxiter, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1 && rowptr && col && Ae && Av && b && x && r && z && p && state && ws,
              "esr_pcg_init: null argument");
  hipStream_t s = as_stream(stream);
  PcgState h{};
  h.rtol = rtol;
  h.atol = atol;
  h.maxiter = maxiter;
  // parameters go in with a plain copy; the kernel fills the rest
  ESR_HIP(hipMemcpyAsync(state, &h, sizeof(h), hipMemcpyHostToDevice, s));
  ESR_HIP(hipStreamSynchronize(s));
  ESR_FLOAT_DISPATCH(dtype, (k_pcg_init<T><<<grid_of(n), kB, 0, s>>>(
                                n, rowptr, col, (const T *)Ae, (const T *)Av, (const T *)diag,
                                (const T *)b, (const T *)x, (T *)r, (T *)z, (T *)p,
                                (PcgState *)state, (double *)ws)));
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

ESR_API int esr_pcg_steps(int dtype, int64_t n, const int32_t *rowptr, const int32_t *col,
                          const void *Ae, const void *Av, const void *diag, void *x, void *r,
                          void *z, void *p, void *Ap, void *state, void *ws, int64_t count,
                          void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1 && count >= 0 && rowptr && col && Ae && Av && x && r && z && p && Ap &&
                  state && ws,
              "esr_pcg_steps: null argument");
  hipStream_t s = as_stream(stream);
  const unsigned g = grid_of(n);
  for (int64_t it = 0; it < count; ++it) {
    ESR_FLOAT_DISPATCH(dtype, {
      k_pcg_spmv<T><<<g, kB, 0, s>>>(n, rowptr, col, (const T *)Ae, (const T *)Av,
                                     (const T *)p, (T *)Ap, (PcgState *)state, (double *)ws);
      k_pcg_xr<T><<<g, kB, 0, s>>>(n, (const T *)diag, (T *)x, (T *)r, (T *)z,
                                   (const T *)p, (const T *)Ap, (PcgState *)state,
                                   (double *)ws);
      k_pcg_p<T><<<g, kB, 0, s>>>(n, (const T *)z, (T *)p, (PcgState *)state);
    });
    ESR_LAUNCHED();
  }
  return ESR_OK;
  ESR_GUARD_END
}

ESR_API int esr_pcg_read(const void *state, double *scalars, int64_t *ints, void *stream) {
  ESR_REQUIRE(state && scalars && ints, "esr_pcg_read: null argument");
  PcgState h;
  hipStream_t s = as_stream(stream);
  ESR_HIP(hipMemcpyAsync(&h, state, sizeof(h), hipMemcpyDeviceToHost, s));
  ESR_HIP(hipStreamSynchronize(s));
  scalars[0] = h.rnorm;
  scalars[1] = h.bnorm;
  scalars[2] = h.rz_cur;
  scalars[3] = h.tol;
  scalars[4] = h.alpha;
  scalars[5] = h.beta;
  ints[0] = h.iters;
  ints[1] = h.done;
  ints[2] = h.k_cur;
  return ESR_OK;
}

ESR_API int esr_gmres_workspace(int64_t n, int64_t restart, int64_t *state_bytes,
                                int64_t *ws_bytes) {
  ESR_REQUIRE(n >= 1 && restart >= 1 && state_bytes && ws_bytes, "esr_gmres_workspace: bad");
  ESR_REQUIRE(restart <= kGmresMaxRestart, "esr_gmres_workspace: restart %lld > %d",
              (long long)restart, kGmresMaxRestart);
  *state_bytes = (int64_t)sizeof(GmresState);
  *ws_bytes = (int64_t)grid_of(n) * (int64_t)sizeof(double);
  return ESR_OK;
}

ESR_API int esr_gmres_begin(int dtype, int64_t n, const void *b, void *state, void *ws,
                            void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1 && b && state && ws, "esr_gmres_begin: null argument");
  hipStream_t s = as_stream(stream);
  GmresState h{};
  ESR_HIP(hipMemcpyAsync(state, &h, sizeof(h), hipMemcpyHostToDevice, s));
  ESR_HIP(hipStreamSynchronize(s));
  ESR_FLOAT_DISPATCH(dtype, (k_gm_bnorm<T><<<grid_of(n), kB, 0, s>>>(
                                n, (const T *)b, (GmresState *)state, (double *)ws)));
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

ESR_API int esr_gmres_residual(int dtype, int64_t n, const int32_t *rowptr, const int32_t *col,
                               const void *Ae, const void *Av, const void *b, const void *x,
                               void *r, void *state, void *ws, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1 && rowptr && col && Ae && Av && b && x && r && state && ws,
              "esr_gmres_residual: null argument");
  hipStream_t s = as_stream(stream);
  ESR_FLOAT_DISPATCH(dtype, (k_gm_residual<T><<<grid_of(n), kB, 0, s>>>(
                                n, rowptr, col, (const T *)Ae, (const T *)Av, (const T *)b,
                                (const T *)x, (T *)r, (GmresState *)state, (double *)ws)));
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

// One Arnoldi cycle: V (restart x n, one basis vector per row), w0/w1 two
// n-vectors used alternately; every kernel is a no-op after a breakdown.
ESR_API int esr_gmres_cycle(int dtype, int64_t n, int64_t restart, const int32_t *rowptr,
                            const int32_t *col, const void *Ae, const void *Av,
                            const void *diag, const void *r, void *V, void *w0, void *w1,
                            void *state, void *ws, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1 && restart >= 1 && restart <= kGmresMaxRestart,
              "esr_gmres_cycle: bad sizes");
  ESR_REQUIRE(rowptr && col && Ae && Av && r && V && w0 && w1 && state && ws,
              "esr_gmres_cycle: null argument");
  hipStream_t s = as_stream(stream);
  const unsigned g = grid_of(n);
  ESR_FLOAT_DISPATCH(dtype, {
    T *Vt = (T *)V;
    T *w[2] = {(T *)w0, (T *)w1};
    GmresState *st = (GmresState *)state;
    for (int j = 0; j < (int)restart; ++j) {
      const T *src = j == 0 ? (const T *)r : w[(j - 1) & 1];
      T *wj = w[j & 1];
      k_gm_spmv<T><<<g, kB, 0, s>>>(n, j, rowptr, col, (const T *)Ae, (const T *)Av,
                                    (const T *)diag, src, Vt + (int64_t)j * n, wj, st);
      for (int i = 0; i <= j; ++i)
        k_gm_mgs<T><<<g, kB, 0, s>>>(n, i, j, i > 0 ? Vt + (int64_t)(i - 1) * n : Vt,
                                     Vt + (int64_t)i * n, wj, st, (double *)ws);
      k_gm_norm<T><<<g, kB, 0, s>>>(n, j, Vt + (int64_t)j * n, wj, st, (double *)ws);
    }
  });
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

ESR_API int esr_gmres_update_x(int dtype, int64_t n, int64_t jcount, const void *V,
                               const void *y, const void *diag, void *x, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1 && jcount >= 0 && jcount <= kGmresMaxRestart && V && y && x,
              "esr_gmres_update_x: bad argument");
  hipStream_t s = as_stream(stream);
  ESR_FLOAT_DISPATCH(dtype, (k_gm_update_x<T><<<grid_of(n), kB, 0, s>>>(
                                n, (int)jcount, (const T *)V, (const T *)y,
                                (const T *)diag, (T *)x)));
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

// Host copy of the cycle's results: scalars = {rnorm, bnorm},
// ints = {jend, broken}, H = (restart+1) x restart row-major.
ESR_API int esr_gmres_read(const void *state, int64_t restart, double *scalars, int64_t *ints,
                           double *H, void *stream) {
  ESR_REQUIRE(state && scalars && ints && restart >= 1 && restart <= kGmresMaxRestart,
              "esr_gmres_read: bad argument");
  GmresState h;
  hipStream_t s = as_stream(stream);
  ESR_HIP(hipMemcpyAsync(&h, state, sizeof(h), hipMemcpyDeviceToHost, s));
  ESR_HIP(hipStreamSynchronize(s));
  scalars[0] = h.rnorm;
  scalars[1] = h.bnorm;
  ints[0] = h.jend;
  ints[1] = h.broken;
  if (H)
    for (int64_t i = 0; i <= restart; ++i)
      for (int64_t j = 0; j < restart; ++j) H[i * restart + j] = h.H[i * kGmresMaxRestart + j];
  return ESR_OK;
}

ESR_API int esr_ir_workspace(int64_t n, int64_t *state_bytes, int64_t *ws_bytes) {
  ESR_REQUIRE(n >= 1 && state_bytes && ws_bytes, "esr_ir_workspace: bad args");
  *state_bytes = (int64_t)sizeof(IrState);
  *ws_bytes = 2 * (int64_t)grid_of(n) * (int64_t)sizeof(double);
  return ESR_OK;
}

ESR_API int esr_ir_residual(int64_t n, const int32_t *rowptr, const int32_t *col, const double *Ae,
                            const double *Av, const double *b, const double *x, float *r32,
                            void *state, void *ws, int first, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1 && rowptr && col && Ae && Av && b && x && r32 && state && ws,
              "esr_ir_residual: null argument");
  hipStream_t s = as_stream(stream);
  if (first) {
    IrState h{};
    ESR_HIP(hipMemcpyAsync(state, &h, sizeof(h), hipMemcpyHostToDevice, s));
    ESR_HIP(hipStreamSynchronize(s));
  }
  k_ir_residual<<<grid_of(n), kB, 0, s>>>(n, rowptr, col, Ae, Av, b, x, r32, (IrState *)state,
                                          (double *)ws, first);
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

ESR_API int esr_ir_update(int64_t n, double *x, float *d, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1 && x && d, "esr_ir_update: null argument");
  k_ir_update<<<grid_of(n), kB, 0, as_stream(stream)>>>(n, x, d);
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

ESR_API int esr_ir_read(const void *state, double *scalars, void *stream) {
  ESR_REQUIRE(state && scalars, "esr_ir_read: null argument");
  IrState h;
  hipStream_t s = as_stream(stream);
  ESR_HIP(hipMemcpyAsync(&h, state, sizeof(h), hipMemcpyDeviceToHost, s));
  ESR_HIP(hipStreamSynchronize(s));
  scalars[0] = h.rnorm;
  scalars[1] = h.bnorm;
  return ESR_OK;
}
