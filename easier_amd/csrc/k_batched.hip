// Batched Jacobi/identity-preconditioned CG over many independent Linsys
// instances (easier/numeric/solver/cg.py:12-137 run on each instance).
//
// Instances are concatenated: global rows [row0_k, row1_k) belong to
// instance k, column indices are global. Instance k is served by its own
// nb_k workgroups ("tiles" first_k .. first_k + nb_k - 1), which stride over
// its rows, so a workgroup never mixes instances and returns at once when its
// instance has converged (per-instance convergence masks). Dot products are
// reduced per instance without atomics or fences: every tile stores one
// partial, and each tile of the next kernel sums its instance's nb_k partials
// in a fixed order (esr_krylov.h), so every instance's scalars are
// deterministic and independent of the other instances.
#include "esr_common.h"
#include "esr_krylov.h"

namespace esr {

struct BatchState {
  double rz_cur, rz_next, pAp, alpha, rr, rz_new, rnorm, bnorm, tol, rtol, atol, pad;
  int64_t k_cur, k_next, maxiter, done, iters, pad2;
};

namespace {

struct Tiles {
  const int32_t *inst;   // instance of each tile
  const int32_t *row0;   // first global row of the tile's instance
  const int32_t *row1;   // one past its last row
  const int32_t *first;  // first tile of each instance
  const int32_t *ntiles; // tiles of each instance
};

// Rows of tile t: its instance's rows, strided over the instance's tiles.
#define ESR_TILE_ROWS(i, tl, t, k)                                                  \
  for (int64_t i = (int64_t)(tl).row0[t] + (int64_t)((t) - (tl).first[k]) * kRB +   \
                   threadIdx.x,                                                     \
               i##_end = (tl).row1[t], i##_step = (int64_t)(tl).ntiles[k] * kRB;    \
       i < i##_end; i += i##_step)

// Init (cg.py:25-29): r = b - A x; z = M r; p = z; partials of r.z, b.b.
__global__ __launch_bounds__(kRB) void k_b_init(
    Tiles tl, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const double *__restrict__ Ae, const double *__restrict__ Av, const double *__restrict__ diag,
    const double *__restrict__ bvec, const double *__restrict__ x, double *__restrict__ r,
    double *__restrict__ z, double *__restrict__ p, double *__restrict__ part) {
  const int t = blockIdx.x;
  const int k = tl.inst[t];
  double acc[2] = {0.0, 0.0};
  ESR_TILE_ROWS(i, tl, t, k) {
    const double ri = bvec[i] - linsys_row<double>(i, rowptr, col, Ae, Av,
                                                   [&](int64_t c) { return x[c]; });
    const double zi = diag ? ri / diag[i] : ri;
    r[i] = ri;
    z[i] = zi;
    p[i] = zi;
    acc[0] += ri * zi;
    acc[1] += bvec[i] * bvec[i];
  }
  store_partials<2>(acc, part);
}

// Per instance: rz, bnorm, tol (cg.py:112-115), counters.
__global__ __launch_bounds__(kRB) void k_b_init_fin(Tiles tl, const double *__restrict__ part,
                                                   BatchState *st) {
  const int t = blockIdx.x;
  const int k = tl.inst[t];
  if (t != tl.first[k]) return;
  double tot[2];
  sum_partials<2>(part, gridDim.x, tl.first[k], tl.ntiles[k], tot);
  if (threadIdx.x == 0) {
    BatchState *s = st + k;
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

// Step part 1 (cg.py:44-47): Ap = A p; partials of p.Ap.
__global__ __launch_bounds__(kRB) void k_b_spmv(
    Tiles tl, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const double *__restrict__ Ae, const double *__restrict__ Av, const double *__restrict__ p,
    double *__restrict__ Ap, BatchState *st, double *__restrict__ part) {
  const int t = blockIdx.x;
  const int k = tl.inst[t];
  BatchState *s = st + k;
  if (s->done) return;
  if (t == tl.first[k] && threadIdx.x == 0) {
    s->k_cur = s->k_next;
    s->rz_cur = s->rz_next;
  }
  double acc[1] = {0.0};
  ESR_TILE_ROWS(i, tl, t, k) {
    const double y = linsys_row<double>(i, rowptr, col, Ae, Av, [&](int64_t c) { return p[c]; });
    Ap[i] = y;
    acc[0] += p[i] * y;
  }
  store_partials<1>(acc, part);
}

// (The product on a padded row-major ELL copy of the concatenated operator
// -- 8 slots per row, no row-pointer loads, as the multi-right-hand-side
// solver reads it -- gave the same bits but took 1.54 s against 1.51 s for
// the 256 meshes of sides 32-512: with a row per thread the padding's bytes
// cost more than the row pointers.)

// Step part 2 (cg.py:47-52) + first half of Update: alpha, x, r, z;
// partials of r.r and r.z.
__global__ __launch_bounds__(kRB) void k_b_xr(Tiles tl, const double *__restrict__ diag,
                                             double *__restrict__ x, double *__restrict__ r,
                                             double *__restrict__ z, const double *__restrict__ p,
                                             const double *__restrict__ Ap, BatchState *st,
                                             const double *__restrict__ part_in,
                                             double *__restrict__ part_out) {
  const int t = blockIdx.x;
  const int k = tl.inst[t];
  BatchState *s = st + k;
  if (s->done) return;
  double pAp[1];
  sum_partials<1>(part_in, gridDim.x, tl.first[k], tl.ntiles[k], pAp);
  const double alpha = s->rz_cur / pAp[0];
  if (t == tl.first[k] && threadIdx.x == 0) {
    s->pAp = pAp[0];
    s->alpha = alpha;
  }
  double acc[2] = {0.0, 0.0};
  ESR_TILE_ROWS(i, tl, t, k) {
    const double ri = r[i] - alpha * Ap[i];
    x[i] = x[i] + alpha * p[i];
    r[i] = ri;
    const double zi = diag ? ri / diag[i] : ri;
    z[i] = zi;
    acc[0] += ri * ri;
    acc[1] += ri * zi;
  }
  store_partials<2>(acc, part_out);
}

// Loop test (cg.py:126-129) and Update (cg.py:67-71) per instance. The
// instance's first tile sets `done` in this kernel while its other tiles read
// it, so both accesses are atomic; every tile derives the same exit anyway.
__global__ __launch_bounds__(kRB) void k_b_p(Tiles tl, const double *__restrict__ z,
                                            double *__restrict__ p, BatchState *st,
                                            const double *__restrict__ part_in,
                                            int32_t *active) {
  const int t = blockIdx.x;
  const int k = tl.inst[t];
  BatchState *s = st + k;
  if (__hip_atomic_load(&s->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  double sums[2];
  sum_partials<2>(part_in, gridDim.x, tl.first[k], tl.ntiles[k], sums);
  const double rnorm = cr_sqrt(sums[0]);
  const int64_t kk = s->k_cur;
  const bool stop = ((rnorm == rnorm) && rnorm <= s->tol) ||
                    (s->maxiter >= 0 && kk >= s->maxiter) || (rnorm != rnorm && s->maxiter < 0);
  const bool lead = t == tl.first[k] && threadIdx.x == 0;
  if (stop) {
    if (lead) {
      s->rr = sums[0];
      s->rz_new = sums[1];
      s->rnorm = rnorm;
      s->iters = kk;
      __hip_atomic_store(&s->done, (int64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      atomicSub(active, 1);
    }
    return;
  }
  const double beta = sums[1] / s->rz_cur;
  ESR_TILE_ROWS(i, tl, t, k) p[i] = z[i] + beta * p[i];
  if (lead) {
    s->rr = sums[0];
    s->rz_new = sums[1];
    s->rnorm = rnorm;
    s->k_next = kk + 1;
    s->rz_next = sums[1];
  }
}

} // namespace
} // namespace esr

using namespace esr;

ESR_API int esr_bcg_state_bytes(int64_t *bytes) {
  ESR_REQUIRE(bytes, "esr_bcg_state_bytes: null");
  *bytes = (int64_t)sizeof(BatchState);
  return ESR_OK;
}

// tiles: inst/row0/row1 (ntiles) and first/ntiles per instance, device int32
// arrays; row0/row1 are the tile's instance's row range. `part` holds
// 3 * ntiles doubles. R (1, 2, 4 or 8) is the rows-per-thread target the
// tiling was built for.
ESR_API int esr_bcg_init(int R, int64_t ntiles, const int32_t *tinst, const int32_t *trow0,
                         const int32_t *trow1, const int32_t *ifirst, const int32_t *intiles,
                         const int32_t *rowptr, const int32_t *col, const double *Ae,
                         const double *Av, const double *diag, const double *b, double *x,
                         double *r, double *z, double *p, void *state, double *part,
                         void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(R == 1 || R == 2 || R == 4 || R == 8, "esr_bcg_init: R must be 1, 2, 4 or 8");
  ESR_REQUIRE(ntiles >= 1 && ntiles < (1ll << 31) && tinst && trow0 && trow1 && ifirst &&
                  intiles && rowptr && col && Ae && Av && b && x && r && z && p && state && part,
              "esr_bcg_init: null argument");
  Tiles tl{tinst, trow0, trow1, ifirst, intiles};
  hipStream_t s = as_stream(stream);
  k_b_init<<<(unsigned)ntiles, kRB, 0, s>>>(tl, rowptr, col, Ae, Av, diag, b, x, r, z, p,
                                            part + ntiles);
  k_b_init_fin<<<(unsigned)ntiles, kRB, 0, s>>>(tl, part + ntiles, (BatchState *)state);
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
  ESR_REQUIRE(R == 1 || R == 2 || R == 4 || R == 8, "esr_bcg_steps: R must be 1, 2, 4 or 8");
  ESR_REQUIRE(ntiles >= 1 && ntiles < (1ll << 31) && count >= 0 && active && part,
              "esr_bcg_steps: bad argument");
  Tiles tl{tinst, trow0, trow1, ifirst, intiles};
  hipStream_t s = as_stream(stream);
  BatchState *st = (BatchState *)state;
  for (int64_t it = 0; it < count; ++it) {
    k_b_spmv<<<(unsigned)ntiles, kRB, 0, s>>>(tl, rowptr, col, Ae, Av, p, Ap, st, part);
    k_b_xr<<<(unsigned)ntiles, kRB, 0, s>>>(tl, diag, x, r, z, p, Ap, st, part,
                                            part + ntiles);
    k_b_p<<<(unsigned)ntiles, kRB, 0, s>>>(tl, z, p, st, part + ntiles, active);
    ESR_LAUNCHED();
  }
  return ESR_OK;
  ESR_GUARD_END
}
