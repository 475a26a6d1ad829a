// Flexible conjugate gradients (Polak-Ribiere direction update) for a
// preconditioner that changes from step to step (the learned network).
// No reference counterpart: the reference's CG (numeric/solver/cg.py:56-71)
// uses the Fletcher-Reeves beta = r1.z1 / r.z, which assumes a fixed M.
//
//   q = A p; alpha = (r.z) / (p.q); r_old = r; r -= alpha q; x += alpha p
//   z = M(r); beta = z.(r - r_old) / (r_old.z_old); p = z + beta p
//
// Scalars stay on the device (FcgState); dot products are deterministic
// (last-workgroup reductions, float64).
#include "esr_common.h"

namespace esr {
struct FcgState {
  double rz;      // r.z of the current direction
  double pq;
  double alpha;
  double rr;      // ||r||^2 (recursive)
  double zr;      // z.r  (new z, new r)
  double zro;     // z.r_old
  double beta;
  double rr_true; // ||b - A x||^2
  double bb;
  uint32_t counter[4];
};

namespace {
constexpr int kB = 256, kR = 4, kT = kB * kR;

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

template <int NV>
__device__ bool sum_last(double (&v)[NV], double *part, uint32_t *counter, int64_t nb, int64_t b,
                         double (&out)[NV]) {
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
      __hip_atomic_store(part + k * nb + b, s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __threadfence();
    last = (atomicAdd(counter, 1u) == (uint32_t)(nb - 1));
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < nb; i += kB)
      s += __hip_atomic_load(part + k * nb + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s = wsum(s);
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

template <typename Src>
__device__ __forceinline__ double spmv(int64_t r, const int32_t *__restrict__ rowptr,
                                       const int32_t *__restrict__ col,
                                       const double *__restrict__ Ae,
                                       const double *__restrict__ Av, Src src) {
  double s = 0.0;
  for (int32_t j = rowptr[r]; j < rowptr[r + 1]; ++j) s = s + src(col[j]) * Ae[j];
  return s + src(r) * Av[r];
}

#define ROWS_OF_BLOCK                                                          \
  const int64_t b = xcd_remap(blockIdx.x, gridDim.x);                          \
  for (int k_ = 0; k_ < kR; ++k_) {                                            \
    const int64_t i = b * kT + k_ * kB + threadIdx.x;                          \
    if (i >= n) break;

__global__ __launch_bounds__(kB) void k_fcg_spmv(int64_t n, const int32_t *rowptr, const int32_t *col,
                                                const double *Ae, const double *Av,
                                                const double *__restrict__ p, double *__restrict__ q,
                                                FcgState *st, double *part) {
  double acc[1] = {0.0};
  ROWS_OF_BLOCK
    const double y = spmv(i, rowptr, col, Ae, Av, [&](int64_t c) { return p[c]; });
    q[i] = y;
    acc[0] += p[i] * y;
  }
  double tot[1];
  if (sum_last<1>(acc, part, &st->counter[0], gridDim.x, b, tot) && threadIdx.x == 0) {
    st->pq = tot[0];
    st->alpha = st->rz / tot[0];
  }
}

__global__ __launch_bounds__(kB) void k_fcg_update(int64_t n, double *__restrict__ x,
                                                  double *__restrict__ r, double *__restrict__ ro,
                                                  const double *__restrict__ p,
                                                  const double *__restrict__ q, FcgState *st,
                                                  double *part) {
  const double alpha = st->alpha;
  double acc[1] = {0.0};
  ROWS_OF_BLOCK
    const double ri = r[i];
    ro[i] = ri;
    const double rn = ri - alpha * q[i];
    r[i] = rn;
    x[i] = x[i] + alpha * p[i];
    acc[0] += rn * rn;
  }
  double tot[1];
  if (sum_last<1>(acc, part, &st->counter[1], gridDim.x, b, tot) && threadIdx.x == 0)
    st->rr = tot[0];
}

// zr = z.r, zro = z.r_old; beta = (zr - zro) / rz; then rz <- zr
__global__ __launch_bounds__(kB) void k_fcg_dots(int64_t n, const double *__restrict__ z,
                                                const double *__restrict__ r,
                                                const double *__restrict__ ro, int restart,
                                                FcgState *st, double *part) {
  double acc[2] = {0.0, 0.0};
  ROWS_OF_BLOCK
    acc[0] += z[i] * r[i];
    acc[1] += z[i] * ro[i];
  }
  double tot[2];
  if (sum_last<2>(acc, part, &st->counter[2], gridDim.x, b, tot) && threadIdx.x == 0) {
    st->zr = tot[0];
    st->zro = tot[1];
    st->beta = restart ? 0.0 : (tot[0] - tot[1]) / st->rz;
    st->rz = tot[0];
  }
}

__global__ __launch_bounds__(kB) void k_fcg_dir(int64_t n, const double *__restrict__ z,
                                               double *__restrict__ p, const FcgState *st) {
  const double beta = st->beta;
  ROWS_OF_BLOCK
    p[i] = z[i] + beta * p[i];
  }
}

// r = b - A x (float64) and ||r||^2, ||b||^2
__global__ __launch_bounds__(kB) void k_fcg_residual(int64_t n, const int32_t *rowptr,
                                                    const int32_t *col, const double *Ae,
                                                    const double *Av, const double *__restrict__ bv,
                                                    const double *__restrict__ x,
                                                    double *__restrict__ r, FcgState *st,
                                                    double *part) {
  double acc[2] = {0.0, 0.0};
  ROWS_OF_BLOCK
    const double ri = bv[i] - spmv(i, rowptr, col, Ae, Av, [&](int64_t c) { return x[c]; });
    r[i] = ri;
    acc[0] += ri * ri;
    acc[1] += bv[i] * bv[i];
  }
  double tot[2];
  if (sum_last<2>(acc, part, &st->counter[3], gridDim.x, b, tot) && threadIdx.x == 0) {
    st->rr_true = tot[0];
    st->rr = tot[0];
    st->bb = tot[1];
  }
}

// z = r / d (the Jacobi fallback)
__global__ __launch_bounds__(kB) void k_fcg_jacobi(int64_t n, const double *__restrict__ r,
                                                  const double *__restrict__ d,
                                                  double *__restrict__ z) {
  ROWS_OF_BLOCK
    z[i] = r[i] / d[i];
  }
}
#undef ROWS_OF_BLOCK

unsigned grid(int64_t n) { return (unsigned)ceil_div(n, kT); }
} // namespace
} // namespace esr

using namespace esr;

ESR_API int esr_fcg_workspace(int64_t n, int64_t *state_bytes, int64_t *ws_bytes) {
  ESR_REQUIRE(n >= 1 && state_bytes && ws_bytes, "esr_fcg_workspace: bad args");
  *state_bytes = (int64_t)sizeof(FcgState);
  *ws_bytes = 2 * (int64_t)grid(n) * (int64_t)sizeof(double);
  return ESR_OK;
}

// op codes: 0 spmv(p -> q), 1 update(x, r, ro, p, q), 2 dots(z, r, ro; a0 = restart),
// 3 direction(z, p), 4 residual(b, x -> r), 5 jacobi(r -> z with d)
ESR_API int esr_fcg_op(int op, int64_t n, const int32_t *rowptr, const int32_t *col,
                       const double *Ae, const double *Av, const double *d, const double *b,
                       double *x, double *r, double *ro, double *p, double *q, double *z,
                       int a0, void *state, void *ws, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1 && state && ws, "esr_fcg_op: null argument");
  hipStream_t s = as_stream(stream);
  FcgState *st = (FcgState *)state;
  double *part = (double *)ws;
  switch (op) {
  case 0:
    ESR_REQUIRE(rowptr && col && Ae && Av && p && q, "esr_fcg_op spmv: null");
    k_fcg_spmv<<<grid(n), kB, 0, s>>>(n, rowptr, col, Ae, Av, p, q, st, part);
    break;
  case 1:
    ESR_REQUIRE(x && r && ro && p && q, "esr_fcg_op update: null");
    k_fcg_update<<<grid(n), kB, 0, s>>>(n, x, r, ro, p, q, st, part);
    break;
  case 2:
    ESR_REQUIRE(z && r && ro, "esr_fcg_op dots: null");
    k_fcg_dots<<<grid(n), kB, 0, s>>>(n, z, r, ro, a0, st, part);
    break;
  case 3:
    ESR_REQUIRE(z && p, "esr_fcg_op direction: null");
    k_fcg_dir<<<grid(n), kB, 0, s>>>(n, z, p, st);
    break;
  case 4:
    ESR_REQUIRE(rowptr && col && Ae && Av && b && x && r, "esr_fcg_op residual: null");
    k_fcg_residual<<<grid(n), kB, 0, s>>>(n, rowptr, col, Ae, Av, b, x, r, st, part);
    break;
  case 5:
    ESR_REQUIRE(r && d && z, "esr_fcg_op jacobi: null");
    k_fcg_jacobi<<<grid(n), kB, 0, s>>>(n, r, d, z);
    break;
  default:
    set_error("esr_fcg_op: unknown op %d", op);
    return ESR_ERR_INVALID;
  }
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

ESR_API int esr_fcg_read(const void *state, double *scalars, void *stream) {
  ESR_REQUIRE(state && scalars, "esr_fcg_read: null");
  FcgState h;
  hipStream_t s = as_stream(stream);
  ESR_HIP(hipMemcpyAsync(&h, state, sizeof(h), hipMemcpyDeviceToHost, s));
  ESR_HIP(hipStreamSynchronize(s));
  const double v[9] = {h.rz, h.pq, h.alpha, h.rr, h.zr, h.zro, h.beta, h.rr_true, h.bb};
  for (int k = 0; k < 9; ++k) scalars[k] = v[k];
  return ESR_OK;
}
