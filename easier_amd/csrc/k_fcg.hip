// Flexible conjugate gradients (Polak-Ribiere direction update) for a
// preconditioner that changes from step to step (the learned network).
// No reference counterpart: the reference's CG (numeric/solver/cg.py:56-71)
// uses the Fletcher-Reeves beta = r1.z1 / r.z, which assumes a fixed M.
//
//   q = A p; alpha = (r.z) / (p.q); r_old = r; r -= alpha q; x += alpha p
//   z = M(r); beta = z.(r - r_old) / (r_old.z_old); p = z + beta p
//
// Scalars stay on the device (FcgState). Each reducing operation is a
// grid-stride kernel storing one partial per workgroup and a one-workgroup
// kernel summing them in a fixed order (esr_krylov.h): deterministic, no
// atomics, no fences.
#include "esr_common.h"
#include "esr_krylov.h"

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
  double tol2;    // the loop exits once rr <= tol2 ...
  int64_t maxit;  // ... or after maxit steps (< 0: no limit) ...
  int64_t it;     // ... steps taken since esr_fcg_set
  int64_t done;   // 1: the step operations are no-ops
  int64_t nbad;   // directions with z.r <= 0 or not finite since esr_fcg_set
  int64_t use_jac; // this step's z was rejected: the direction uses r / d
};

namespace {

template <typename Op>
__global__ __launch_bounds__(kRB) void k_fcg_spmv(int64_t n, const Op A,
                                                 const double *__restrict__ p,
                                                 double *__restrict__ q, const FcgState *st,
                                                 double *part) {
  if (st->done) return;
  double acc[1] = {0.0};
  ESR_GRID_ROWS(i, n) {
    const double y = A.row(i, [&](int64_t c) { return p[c]; });
    q[i] = y;
    acc[0] += p[i] * y;
  }
  store_partials<1>(acc, part);
}

__global__ __launch_bounds__(kRB) void k_fcg_update(int64_t n, double *__restrict__ x,
                                                   double *__restrict__ r,
                                                   double *__restrict__ ro,
                                                   const double *__restrict__ p,
                                                   const double *__restrict__ q,
                                                   const FcgState *st, double *part) {
  if (st->done) return;
  const double alpha = st->alpha;
  double acc[1] = {0.0};
  ESR_GRID_ROWS(i, n) {
    const double ri = r[i];
    ro[i] = ri;
    const double rn = ri - alpha * q[i];
    r[i] = rn;
    x[i] = x[i] + alpha * p[i];
    acc[0] += rn * rn;
  }
  store_partials<1>(acc, part);
}

// z.r, z.r_old and the same for the Jacobi vector j = r / d, which replaces
// z for this step when z is not a descent direction (z.r <= 0 or not finite).
__global__ __launch_bounds__(kRB) void k_fcg_dots(int64_t n, const double *__restrict__ z,
                                                 const double *__restrict__ r,
                                                 const double *__restrict__ ro,
                                                 const double *__restrict__ d,
                                                 const FcgState *st, double *part) {
  if (st->done) return;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  ESR_GRID_ROWS(i, n) {
    const double ji = r[i] / d[i];
    acc[0] += z[i] * r[i];
    acc[1] += z[i] * ro[i];
    acc[2] += ji * r[i];
    acc[3] += ji * ro[i];
  }
  store_partials<4>(acc, part);
}

// p = s + beta p, or p = s on a restart; s = z, or r / d for a step whose z
// was rejected.
__global__ __launch_bounds__(kRB) void k_fcg_dir(int64_t n, const double *__restrict__ z,
                                                const double *__restrict__ r,
                                                const double *__restrict__ d,
                                                double *__restrict__ p, const FcgState *st,
                                                int restart) {
  if (st->done) return;
  const double beta = st->beta;
  const bool jac = st->use_jac != 0;
  ESR_GRID_ROWS(i, n) {
    const double si = jac ? r[i] / d[i] : z[i];
    p[i] = restart ? si : si + beta * p[i];
  }
}

// r = b - A x (float64); partials of ||r||^2 and ||b||^2
template <typename Op>
__global__ __launch_bounds__(kRB) void k_fcg_residual(
    int64_t n, const Op A, const double *__restrict__ bv,
    const double *__restrict__ x, double *__restrict__ r, double *part) {
  double acc[2] = {0.0, 0.0};
  ESR_GRID_ROWS(i, n) {
    const double ri = bv[i] - A.row(i, [&](int64_t c) { return x[c]; });
    r[i] = ri;
    acc[0] += ri * ri;
    acc[1] += bv[i] * bv[i];
  }
  store_partials<2>(acc, part);
}

// z = r / d (the Jacobi fallback)
__global__ __launch_bounds__(kRB) void k_fcg_jacobi(int64_t n, const double *__restrict__ r,
                                                   const double *__restrict__ d,
                                                   double *__restrict__ z, const FcgState *st) {
  if (st->done) return;
  ESR_GRID_ROWS(i, n) z[i] = r[i] / d[i];
}

// The scalars of a reducing operation, from its partials (one workgroup):
// spmv: pq, alpha = rz / pq; update: rr, one more step, exit test;
// dots: zr, zro, beta, rz <- zr (beta = 0 on a restart), non-descent count;
// residual: rr_true = rr, bb. The step operations do nothing once `done`.
__global__ __launch_bounds__(kRB) void k_fcg_fin(int op, int64_t g, const double *part,
                                                FcgState *st, int restart) {
  if (op != 4 && st->done) return;
  double tot[4] = {0.0, 0.0, 0.0, 0.0};
  if (op == 0 || op == 1) {
    double v[1];
    sum_partials<1>(part, g, 0, g, v);
    tot[0] = v[0];
  } else if (op == 4) {
    double v[2];
    sum_partials<2>(part, g, 0, g, v);
    tot[0] = v[0];
    tot[1] = v[1];
  } else {
    sum_partials<4>(part, g, 0, g, tot);
  }
  if (threadIdx.x != 0) return;
  switch (op) {
  case 0:
    st->pq = tot[0];
    st->alpha = st->rz / tot[0];
    break;
  case 1: {
    st->rr = tot[0];
    const int64_t it = st->it + 1;
    st->it = it;
    if (!(tot[0] > st->tol2) || (st->maxit >= 0 && it >= st->maxit)) st->done = 1;
    break;
  }
  case 2: {
    const bool bad = !(tot[0] > 0.0) || !isfinite(tot[0]) || !isfinite(tot[1]);
    const double zr = bad ? tot[2] : tot[0], zro = bad ? tot[3] : tot[1];
    st->use_jac = bad ? 1 : 0;
    if (bad) st->nbad += 1;
    st->zr = zr;
    st->zro = zro;
    st->beta = restart ? 0.0 : (zr - zro) / st->rz;
    st->rz = zr;
    break;
  }
  case 4:
    st->rr_true = tot[0];
    st->rr = tot[0];
    st->bb = tot[1];
    break;
  default:
    break;
  }
}

} // namespace
} // namespace esr

using namespace esr;

// Partials: up to four values per workgroup, G = red_grid(n) workgroups.
ESR_API int esr_fcg_workspace(int64_t n, int64_t *state_bytes, int64_t *ws_bytes) {
  ESR_REQUIRE(n >= 1 && state_bytes && ws_bytes, "esr_fcg_workspace: bad args");
  *state_bytes = (int64_t)sizeof(FcgState);
  *ws_bytes = 4 * (int64_t)red_grid(n) * (int64_t)sizeof(double);
  return ESR_OK;
}

namespace {
// op codes: 0 spmv(p -> q), 1 update(x, r, ro, p, q), 2 dots(z, r, ro, d;
// a0 = restart), 3 direction(z or r/d, p; a0 = restart), 4 residual(b, x -> r),
// 5 jacobi(r -> z with d)
template <typename Op>
int fcg_op(int op, int64_t n, const Op &A, bool have_op, const double *d, const double *b,
           double *x, double *r, double *ro, double *p, double *q, double *z, int a0,
           void *state, void *ws, hipStream_t s) {
  ESR_REQUIRE(n >= 1 && state && ws, "esr_fcg_op: null argument");
  FcgState *st = (FcgState *)state;
  double *part = (double *)ws;
  const unsigned g = red_grid(n);
  switch (op) {
  case 0:
    ESR_REQUIRE(have_op && p && q, "esr_fcg_op spmv: null");
    k_fcg_spmv<<<g, kRB, 0, s>>>(n, A, p, q, st, part);
    break;
  case 1:
    ESR_REQUIRE(x && r && ro && p && q, "esr_fcg_op update: null");
    k_fcg_update<<<g, kRB, 0, s>>>(n, x, r, ro, p, q, st, part);
    break;
  case 2:
    ESR_REQUIRE(z && r && ro && d, "esr_fcg_op dots: null");
    k_fcg_dots<<<g, kRB, 0, s>>>(n, z, r, ro, d, st, part);
    break;
  case 3:
    ESR_REQUIRE(z && r && d && p, "esr_fcg_op direction: null");
    k_fcg_dir<<<g, kRB, 0, s>>>(n, z, r, d, p, st, a0);
    break;
  case 4:
    ESR_REQUIRE(have_op && b && x && r, "esr_fcg_op residual: null");
    k_fcg_residual<<<g, kRB, 0, s>>>(n, A, b, x, r, part);
    break;
  case 5:
    ESR_REQUIRE(r && d && z, "esr_fcg_op jacobi: null");
    k_fcg_jacobi<<<g, kRB, 0, s>>>(n, r, d, z, st);
    break;
  default:
    set_error("esr_fcg_op: unknown op %d", op);
    return ESR_ERR_INVALID;
  }
  if (op == 0 || op == 1 || op == 2 || op == 4)
    k_fcg_fin<<<1, kRB, 0, s>>>(op, g, part, st, a0);
  ESR_LAUNCHED();
  return ESR_OK;
}
} // namespace

ESR_API int esr_fcg_op(int op, int64_t n, const int32_t *rowptr, const int32_t *col,
                       const double *Ae, const double *Av, const double *d, const double *b,
                       double *x, double *r, double *ro, double *p, double *q, double *z,
                       int a0, void *state, void *ws, void *stream) {
  ESR_GUARD_BEGIN
  const CsrRows<double> A{rowptr, col, Ae, Av};
  return fcg_op(op, n, A, rowptr && col && Ae && Av, d, b, x, r, ro, p, q, z, a0, state, ws,
                as_stream(stream));
  ESR_GUARD_END
}

// esr_fcg_op with the operator (ops 0 and 4) in its SELL-64 form.
ESR_API int esr_fcg_op_sell(int op, int64_t n, const int64_t *soff, const uint8_t *cnt,
                            const void *colv, int32_t idx16, const double *val,
                            const double *Av, const double *d, const double *b, double *x,
                            double *r, double *ro, double *p, double *q, double *z, int a0,
                            void *state, void *ws, void *stream) {
  ESR_GUARD_BEGIN
  const bool have = soff && cnt && colv && val && Av;
  if (idx16) {
    const SellRows<double, int16_t> A{soff, cnt, (const int16_t *)colv, val, Av};
    return fcg_op(op, n, A, have, d, b, x, r, ro, p, q, z, a0, state, ws, as_stream(stream));
  }
  const SellRows<double, int32_t> A{soff, cnt, (const int32_t *)colv, val, Av};
  return fcg_op(op, n, A, have, d, b, x, r, ro, p, q, z, a0, state, ws, as_stream(stream));
  ESR_GUARD_END
}

// scalars = {rz, pq, alpha, rr, zr, zro, beta, rr_true, bb};
// ints (nullable) = {it, done, nbad}.
ESR_API int esr_fcg_read(const void *state, double *scalars, int64_t *ints, void *stream) {
  ESR_REQUIRE(state && scalars, "esr_fcg_read: null");
  FcgState h;
  hipStream_t s = as_stream(stream);
  ESR_HIP(hipMemcpyAsync(&h, state, sizeof(h), hipMemcpyDeviceToHost, s));
  ESR_HIP(hipStreamSynchronize(s));
  const double v[9] = {h.rz, h.pq, h.alpha, h.rr, h.zr, h.zro, h.beta, h.rr_true, h.bb};
  for (int k = 0; k < 9; ++k) scalars[k] = v[k];
  if (ints) {
    ints[0] = h.it;
    ints[1] = h.done;
    ints[2] = h.nbad;
  }
  return ESR_OK;
}

namespace {
__global__ void k_fcg_set(FcgState *st, double tol2, int64_t maxit) {
  st->tol2 = tol2;
  st->maxit = maxit;
  st->it = 0;
  st->done = 0;
  st->nbad = 0;
  st->use_jac = 0;
}
} // namespace

// Arms the step operations: exit once rr <= tol2 or after maxit steps
// (< 0: no limit); clears the step and non-descent counters.
ESR_API int esr_fcg_set(void *state, double tol2, int64_t maxit, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(state, "esr_fcg_set: null state");
  k_fcg_set<<<1, 1, 0, as_stream(stream)>>>((FcgState *)state, tol2, maxit);
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}
