// Batched Jacobi/identity-preconditioned CG over many independent Linsys
// instances (easier/numeric/solver/cg.py:12-137 run on each instance).
//
// Instances are concatenated: global rows [row0_k, row0_k + n_k) belong to
// instance k, column indices are global. Work is cut into tiles of 256*R rows
// that never straddle an instance, so every workgroup serves one instance and
// can skip the whole tile once its instance has converged (per-instance
// convergence masks). Dot products are reduced per instance: tile partials in
// tile order, finalised by the instance's last tile to arrive (a per-instance
// integer arrival counter), so every instance's scalars are deterministic and
// identical to a solve of that instance alone with the same tiling.
#include "esr_common.h"

namespace esr {

struct BatchState {
  double rz_cur, rz_next, pAp, alpha, rr, rz_new, rnorm, bnorm, tol, rtol, atol, pad;
  int64_t k_cur, k_next, maxiter, done, iters, pad2;
  uint32_t counter[4];
};

namespace {

constexpr int kB = 256;

struct Tiles {
  const int32_t *inst;   // instance of each tile
  const int32_t *row0;   // first global row of the tile
  const int32_t *row1;   // one past the last row
  const int32_t *first;  // first tile of each instance
  const int32_t *ntiles; // tiles of each instance
};

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Per-instance version of the last-block reduction: partials are indexed by
// tile; the instance's tiles are contiguous [first, first + ntiles).
template <int NV>
__device__ bool inst_sum_last(double (&v)[NV], double *part, int64_t ntot, int t,
                              int first, int nt, uint32_t *counter, double (&out)[NV]) {
  __shared__ double sh[NV][kB / 64];
  __shared__ int last;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double s = wsum(v[k]);
    if (lane == 0) sh[k][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int w = 0; w < kB / 64; ++w) s += sh[k][w];
      __hip_atomic_store(part + k * ntot + t, s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __threadfence();
    const uint32_t prev = atomicAdd(counter, 1u);
    last = (prev == (uint32_t)(nt - 1));
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nt; i += kB)
      s += __hip_atomic_load(part + k * ntot + first + i, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
    s = wsum(s);
    __syncthreads();
    if (lane == 0) sh[k][wid] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double tt = 0.0;
      for (int w = 0; w < kB / 64; ++w) tt += sh[k][w];
      out[k] = tt;
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

template <int R>
__global__ __launch_bounds__(kB) void k_b_init(
    Tiles tl, int64_t ntiles, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const double *__restrict__ Ae, const double *__restrict__ Av, const double *__restrict__ diag,
    const double *__restrict__ bvec, const double *__restrict__ x, double *__restrict__ r,
    double *__restrict__ z, double *__restrict__ p, BatchState *st, double *part) {
  const int t = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int k = tl.inst[t];
  const int32_t r0 = tl.row0[t], r1 = tl.row1[t];
  double acc[2] = {0.0, 0.0};
  for (int q = 0; q < R; ++q) {
    const int64_t i = (int64_t)r0 + q * kB + threadIdx.x;
    if (i >= r1) break;
    const double ri = bvec[i] - spmv_row<double>(i, rowptr, col, Ae, Av,
                                                 [&](int64_t c) { return x[c]; });
    const double zi = diag ? ri / diag[i] : ri;
    r[i] = ri;
    z[i] = zi;
    p[i] = zi;
    acc[0] += ri * zi;
    acc[1] += bvec[i] * bvec[i];
  }
  double tot[2];
  BatchState *s = st + k;
  if (inst_sum_last<2>(acc, part, ntiles, t, tl.first[k], tl.ntiles[k], &s->counter[0], tot) &&
      threadIdx.x == 0) {
    s->rz_cur = s->rz_next = tot[0];
    s->bnorm = cr_sqrt(tot[1]);
    double tol = s->rtol * s->bnorm;
    if (s->atol > 0.0 && s->atol > tol) tol = s->atol;
    s->tol = tol;
    s->k_cur = s->k_next = 0;
    s->done = 0;
    s->iters = 0;
  }
}

template <int R>
__global__ __launch_bounds__(kB) void k_b_spmv(
    Tiles tl, int64_t ntiles, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const double *__restrict__ Ae, const double *__restrict__ Av, const double *__restrict__ p,
    double *__restrict__ Ap, BatchState *st, double *part) {
  const int t = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int k = tl.inst[t];
  BatchState *s = st + k;
  if (s->done) return;
  if (t == tl.first[k] && threadIdx.x == 0) {
    s->k_cur = s->k_next;
    s->rz_cur = s->rz_next;
  }
  const int32_t r0 = tl.row0[t], r1 = tl.row1[t];
  double acc[1] = {0.0};
  for (int q = 0; q < R; ++q) {
    const int64_t i = (int64_t)r0 + q * kB + threadIdx.x;
    if (i >= r1) break;
    const double y = spmv_row<double>(i, rowptr, col, Ae, Av, [&](int64_t c) { return p[c]; });
    Ap[i] = y;
    acc[0] += p[i] * y;
  }
  double tot[1];
  if (inst_sum_last<1>(acc, part, ntiles, t, tl.first[k], tl.ntiles[k], &s->counter[1], tot) &&
      threadIdx.x == 0) {
    s->pAp = tot[0];
    s->alpha = s->rz_next / tot[0];
  }
}

template <int R>
__global__ __launch_bounds__(kB) void k_b_xr(Tiles tl, int64_t ntiles,
                                            const double *__restrict__ diag,
                                            double *__restrict__ x, double *__restrict__ r,
                                            double *__restrict__ z, const double *__restrict__ p,
                                            const double *__restrict__ Ap, BatchState *st,
                                            double *part) {
  const int t = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int k = tl.inst[t];
  BatchState *s = st + k;
  if (s->done) return;
  const double alpha = s->alpha;
  const int32_t r0 = tl.row0[t], r1 = tl.row1[t];
  double acc[2] = {0.0, 0.0};
  for (int q = 0; q < R; ++q) {
    const int64_t i = (int64_t)r0 + q * kB + threadIdx.x;
    if (i >= r1) break;
    const double ri = r[i] - alpha * Ap[i];
    x[i] = x[i] + alpha * p[i];
    r[i] = ri;
    const double zi = diag ? ri / diag[i] : ri;
    z[i] = zi;
    acc[0] += ri * ri;
    acc[1] += ri * zi;
  }
  double tot[2];
  if (inst_sum_last<2>(acc, part, ntiles, t, tl.first[k], tl.ntiles[k], &s->counter[2], tot) &&
      threadIdx.x == 0) {
    s->rr = tot[0];
    s->rz_new = tot[1];
    s->rnorm = cr_sqrt(tot[0]);
  }
}

template <int R>
__global__ __launch_bounds__(kB) void k_b_p(Tiles tl, const double *__restrict__ z,
                                           double *__restrict__ p, BatchState *st,
                                           int32_t *active) {
  const int t = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int k = tl.inst[t];
  BatchState *s = st + k;
  if (s->done) return;
  const double rnorm = s->rnorm;
  const int64_t kk = s->k_cur;
  const bool stop = ((rnorm == rnorm) && rnorm <= s->tol) ||
                    (s->maxiter >= 0 && kk >= s->maxiter) || (rnorm != rnorm && s->maxiter < 0);
  if (stop) {
    if (t == tl.first[k] && threadIdx.x == 0) {
      s->done = 1;
      s->iters = kk;
      atomicSub(active, 1);
    }
    return;
  }
  const double beta = s->rz_new / s->rz_cur;
  const int32_t r0 = tl.row0[t], r1 = tl.row1[t];
  for (int q = 0; q < R; ++q) {
    const int64_t i = (int64_t)r0 + q * kB + threadIdx.x;
    if (i >= r1) break;
    p[i] = z[i] + beta * p[i];
  }
  if (t == tl.first[k] && threadIdx.x == 0) {
    s->k_next = kk + 1;
    s->rz_next = s->rz_new;
  }
}

} // namespace
} // namespace esr

using namespace esr;

#define ESR_R_DISPATCH(R, ...)                                                 \
  do {                                                                         \
    switch (R) {                                                               \
    case 1: { constexpr int RR = 1; __VA_ARGS__; } break;                      \
    case 2: { constexpr int RR = 2; __VA_ARGS__; } break;                      \
    case 4: { constexpr int RR = 4; __VA_ARGS__; } break;                      \
    case 8: { constexpr int RR = 8; __VA_ARGS__; } break;                      \
    default: set_error("rows per thread must be 1, 2, 4 or 8"); return ESR_ERR_INVALID; \
    }                                                                          \
  } while (0)

ESR_API int esr_bcg_state_bytes(int64_t *bytes) {
  ESR_REQUIRE(bytes, "esr_bcg_state_bytes: null");
  *bytes = (int64_t)sizeof(BatchState);
  return ESR_OK;
}

// tiles: inst/row0/row1 (ntiles) and first/ntiles per instance (ninst),
// all device int32 arrays built for rows-per-thread R (tile = 256 R rows).
ESR_API int esr_bcg_init(int R, int64_t ntiles, const int32_t *tinst, const int32_t *trow0,
                         const int32_t *trow1, const int32_t *ifirst, const int32_t *intiles,
                         const int32_t *rowptr, const int32_t *col, const double *Ae,
                         const double *Av, const double *diag, const double *b, double *x,
                         double *r, double *z, double *p, void *state, double *part,
                         void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(ntiles >= 1 && tinst && trow0 && trow1 && ifirst && intiles && rowptr && col &&
                  Ae && Av && b && x && r && z && p && state && part,
              "esr_bcg_init: null argument");
  Tiles tl{tinst, trow0, trow1, ifirst, intiles};
  hipStream_t s = as_stream(stream);
  ESR_R_DISPATCH(R, (k_b_init<RR><<<(unsigned)ntiles, kB, 0, s>>>(
                        tl, ntiles, rowptr, col, Ae, Av, diag, b, x, r, z, p,
                        (BatchState *)state, part)));
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

ESR_API int esr_bcg_steps(int R, int64_t ntiles, const int32_t *tinst, const int32_t *trow0,
                          const int32_t *trow1, const int32_t *ifirst, const int32_t *intiles,
                          const int32_t *rowptr, const int32_t *col, const double *Ae,
                          const double *Av, const double *diag, double *x, double *r,
                          double *z, double *p, double *Ap, void *state, double *part,
                          int32_t *active, int64_t count, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(ntiles >= 1 && count >= 0 && active, "esr_bcg_steps: bad argument");
  Tiles tl{tinst, trow0, trow1, ifirst, intiles};
  hipStream_t s = as_stream(stream);
  for (int64_t it = 0; it < count; ++it) {
    ESR_R_DISPATCH(R, {
      k_b_spmv<RR><<<(unsigned)ntiles, kB, 0, s>>>(tl, ntiles, rowptr, col, Ae, Av, p, Ap,
                                                   (BatchState *)state, part);
      k_b_xr<RR><<<(unsigned)ntiles, kB, 0, s>>>(tl, ntiles, diag, x, r, z, p, Ap,
                                                 (BatchState *)state, part);
      k_b_p<RR><<<(unsigned)ntiles, kB, 0, s>>>(tl, z, p, (BatchState *)state, active);
    });
    ESR_LAUNCHED();
  }
  return ESR_OK;
  ESR_GUARD_END
}
