// Jacobi/identity-preconditioned CG for K right-hand sides of one Linsys
// operator (easier/numeric/solver/cg.py:12-137 run on each right-hand side),
// the operator read once per iteration for all of them.
//
// Vectors are stored right-hand-side-minor, v[row * K + k]. A row is served
// by L = (K rounded up to a power of two) consecutive lanes, lane k holding
// right-hand side k, so the row's column indices and stiffness values are
// loaded once (the lanes read the same address) and every gathered
// neighbour row is one contiguous K-value load. Each right-hand side keeps
// its own scalars and convergence flag (BatchState, as the batched solver);
// a converged right-hand side's lanes stop writing. Dot products: one
// partial per (workgroup, right-hand side), summed in a fixed order by a
// one-workgroup-per-right-hand-side kernel that also applies the scalar
// step of cg.py (alpha; loop test, beta), so the row kernels read finished
// scalars. Deterministic, no atomics on floats, no fences.
#include "esr_common.h"
#include "esr_krylov.h"

namespace esr {

struct MrhsState {  // per right-hand side (same layout as the batched solver's)
  double rz_cur, rz_next, pAp, alpha, rr, rz_new, rnorm, bnorm, tol, rtol, atol, beta;
  // xpend: this step's x += alpha p is still to be applied (set by k_m_alpha
  // while the right-hand side runs, cleared by it once the side is done)
  int64_t k_cur, k_next, maxiter, done, iters, xpend;
};

namespace {

constexpr int kMaxK = 64;
constexpr int kMaxGridM = 512;
constexpr int kMaxGridSpmvM = 8192;

struct Lanes {
  int L;        // lanes per row (power of two >= K)
  int rows;     // rows per workgroup = kRB / L
  int k;        // this thread's right-hand side
  int rib;      // this thread's row within the workgroup's row group
};

__device__ __forceinline__ Lanes lanes_of(int K) {
  int L = 1;
  while (L < K) L <<= 1;
  Lanes ln;
  ln.L = L;
  ln.rows = kRB / L;
  ln.k = threadIdx.x % L;
  ln.rib = threadIdx.x / L;
  return ln;
}

#define ESR_MRHS_ROWS(i, n, ln)                                                       \
  for (int64_t i = (int64_t)blockIdx.x * (ln).rows + (ln).rib; i < (n);              \
       i += (int64_t)gridDim.x * (ln).rows)

// Per-(workgroup, right-hand side) sums of NV per-thread values, fixed
// order; part[(v * K + k) * G + blockIdx.x].
template <int NV>
__device__ __forceinline__ void store_rhs_partials(const double (&v)[NV], const Lanes &ln, int K,
                                                   double *part) {
  __shared__ double sh[NV][kRB];
#pragma unroll
  for (int q = 0; q < NV; ++q) sh[q][threadIdx.x] = v[q];
  __syncthreads();
  if (threadIdx.x < K) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double s = 0.0;
      for (int r = 0; r < ln.rows; ++r) s += sh[q][r * ln.L + threadIdx.x];
      part[((int64_t)q * K + threadIdx.x) * gridDim.x + blockIdx.x] = s;
    }
  }
}

// Init (cg.py:25-29) per right-hand side: r = b - A x; z = M r; p = z.
__global__ __launch_bounds__(kRB) void k_m_init(int64_t n, int K, const int32_t *__restrict__ rowptr,
                                               const int32_t *__restrict__ col,
                                               const double *__restrict__ Ae,
                                               const double *__restrict__ Av,
                                               const double *__restrict__ diag,
                                               const double *__restrict__ B,
                                               const double *__restrict__ X,
                                               double *__restrict__ R, double *__restrict__ Z,
                                               double *__restrict__ P, double *__restrict__ part) {
  const Lanes ln = lanes_of(K);
  const bool on = ln.k < K;
  double acc[2] = {0.0, 0.0};
  ESR_MRHS_ROWS(i, n, ln) {
    if (!on) continue;
    const double ax =
        linsys_row<double>(i, rowptr, col, Ae, Av, [&](int64_t c) { return X[c * K + ln.k]; });
    const double b = B[i * K + ln.k];
    const double ri = b - ax;
    const double zi = diag ? ri / diag[i] : ri;
    R[i * K + ln.k] = ri;
    Z[i * K + ln.k] = zi;
    P[i * K + ln.k] = zi;
    acc[0] += ri * zi;
    acc[1] += b * b;
  }
  store_rhs_partials<2>(acc, ln, K, part);
}

// One workgroup per right-hand side: rz, bnorm, tol (cg.py:112-115).
__global__ __launch_bounds__(kRB) void k_m_init_fin(int K, int64_t g, const double *__restrict__ part,
                                                   MrhsState *st, int32_t *active) {
  const int k = blockIdx.x;
  double rz[1], bb[1];
  sum_partials<1>(part, 0, (int64_t)k * g, g, rz);
  sum_partials<1>(part, 0, ((int64_t)K + k) * g, g, bb);
  if (threadIdx.x == 0) {
    MrhsState *s = st + k;
    s->rz_cur = s->rz_next = rz[0];
    s->bnorm = cr_sqrt(bb[0]);
    double tol = s->rtol * s->bnorm;
    if (s->atol > 0.0 && s->atol > tol) tol = s->atol;
    s->tol = tol;
    s->k_cur = s->k_next = 0;
    s->done = 0;
    s->iters = 0;
    s->xpend = 0;
    if (k == 0) *active = K;
  }
}

// Step part 1 (cg.py:44-47): AP = A P; partials of p.Ap. On 64 right-hand
// sides of the 1024 x 1024 mesh this kernel took 0.91 ms of a 1.9 ms step
// on a 512-workgroup grid (1.3 TB/s on its 1.16 GB:
// a wavefront serves one row, 513 rows in turn); on its own grid of up to
// 8,192 workgroups the solve takes 6.03 s instead of 6.48 s (2,048: 6.21 s,
// 32,768: 6.12 s). It fetches 2.2 GB through the L2s for 0.63 GB of
// compulsory reads (neighbour rows fetched by several XCDs), but a
// contiguous chunk of rows per workgroup with the chunks of one XCD
// consecutive did not shorten the solve (6.08 s), nor did two rows per
// wavefront iteration (6.43 s on the 512-workgroup grid).
__global__ __launch_bounds__(kRB) void k_m_spmv(int64_t n, int K, const int32_t *__restrict__ rowptr,
                                               const int32_t *__restrict__ col,
                                               const double *__restrict__ Ae,
                                               const double *__restrict__ Av,
                                               const double *__restrict__ P,
                                               double *__restrict__ AP, const MrhsState *st,
                                               double *__restrict__ part) {
  const Lanes ln = lanes_of(K);
  const bool on = ln.k < K && !st[ln.k < K ? ln.k : 0].done;
  double acc[1] = {0.0};
  ESR_MRHS_ROWS(i, n, ln) {
    if (!on) continue;
    const double y =
        linsys_row<double>(i, rowptr, col, Ae, Av, [&](int64_t c) { return P[c * K + ln.k]; });
    AP[i * K + ln.k] = y;
    acc[0] += P[i * K + ln.k] * y;
  }
  store_rhs_partials<1>(acc, ln, K, part);
}

// The same product on a row-major padded ELL copy of the operator: row i's
// entries at ecol/eval[8 i + k] for k < len[i] (padding: column i, value
// 0, not folded), so a row's column indices and values are two loads that
// do not wait on row pointers, and the next row's are loaded while this
// row's neighbours of P are gathered (64 right-hand sides on 1024^2: 5.51 s
// on CSR, 5.10 s). The row is folded exactly as
// linsys_row folds it (entries left to right from 0, then the diagonal
// term), the rows are visited in the same order, so AP and the partials
// are the CSR kernel's bits.
constexpr int kEllW = ESR_MRHS_ELL_W;
__global__ __launch_bounds__(kRB) void k_m_spmv_ell(int64_t n, int K,
                                                   const uint8_t *__restrict__ len,
                                                   const int32_t *__restrict__ ecol,
                                                   const double *__restrict__ eval,
                                                   const double *__restrict__ Av,
                                                   const double *__restrict__ P,
                                                   double *__restrict__ AP, const MrhsState *st,
                                                   double *__restrict__ part) {
  const Lanes ln = lanes_of(K);
  const bool on = ln.k < K && !st[ln.k < K ? ln.k : 0].done;
  double acc[1] = {0.0};
  // rows in the order of k_m_spmv (one contiguous eighth of the rows per
  // XCD instead, so that a row's neighbours sit in the XCD's own L2, took
  // 5.23 s against 5.10 s for 64 right-hand sides on 1024^2)
  const int64_t stride = (int64_t)gridDim.x * ln.rows;
  int64_t i = (int64_t)blockIdx.x * ln.rows + ln.rib;
  const int64_t iend = n;
  if (on && i < iend) {
    int32_t c[kEllW];
    double v[kEllW];
    int m = len[i];
#pragma unroll
    for (int k = 0; k < kEllW; ++k) {
      c[k] = ecol[i * kEllW + k];
      v[k] = eval[i * kEllW + k];
    }
    for (;;) {
      const int64_t inext = i + stride;
      const bool more = inext < iend;
      // the next row's entries, in flight while this row's gathers are
      int32_t cn[kEllW];
      double vn[kEllW];
      int mn = 0;
      if (more) {
        mn = len[inext];
#pragma unroll
        for (int k = 0; k < kEllW; ++k) {
          cn[k] = ecol[inext * kEllW + k];
          vn[k] = eval[inext * kEllW + k];
        }
      }
      double pv[kEllW];
#pragma unroll
      for (int k = 0; k < kEllW; ++k) pv[k] = P[(int64_t)c[k] * K + ln.k];
      const double pi = P[i * K + ln.k];
      double sum = 0.0;
#pragma unroll
      for (int k = 0; k < kEllW; ++k) {
        const double t = pv[k] * v[k];
        sum = k < m ? sum + t : sum;
      }
      const double y = sum + pi * Av[i];
      AP[i * K + ln.k] = y;
      acc[0] += pi * y;
      if (!more) break;
      i = inext;
      m = mn;
#pragma unroll
      for (int k = 0; k < kEllW; ++k) {
        c[k] = cn[k];
        v[k] = vn[k];
      }
    }
  }
  store_rhs_partials<1>(acc, ln, K, part);
}

// alpha = rz / pAp per right-hand side (one workgroup each); also commits
// the counters of the previous Update.
__global__ __launch_bounds__(kRB) void k_m_alpha(int K, int64_t g, const double *__restrict__ part,
                                                MrhsState *st) {
  const int k = blockIdx.x;
  MrhsState *s = st + k;
  if (s->done) {
    if (threadIdx.x == 0) s->xpend = 0;
    return;
  }
  double tot[1];
  sum_partials<1>(part, 0, (int64_t)k * g, g, tot);
  if (threadIdx.x == 0) {
    s->k_cur = s->k_next;
    s->rz_cur = s->rz_next;
    s->pAp = tot[0];
    s->alpha = s->rz_cur / tot[0];
    s->xpend = 1;
  }
}

// Step part 2 (cg.py:47-52) + first half of Update: r and the partials of
// r.r and r.z with z = M r. x is updated by k_m_p (which reads p anyway) and
// z is recomputed there instead of stored: three vector passes fewer per
// step, the same values (the same operations on the same operands).
__global__ __launch_bounds__(kRB) void k_m_xr(int64_t n, int K, const double *__restrict__ diag,
                                             double *__restrict__ R,
                                             const double *__restrict__ AP, const MrhsState *st,
                                             double *__restrict__ part) {
  const Lanes ln = lanes_of(K);
  const MrhsState *s = st + (ln.k < K ? ln.k : 0);
  const bool on = ln.k < K && !s->done;
  const double alpha = s->alpha;
  double acc[2] = {0.0, 0.0};
  ESR_MRHS_ROWS(i, n, ln) {
    if (!on) continue;
    const int64_t o = i * K + ln.k;
    const double ri = R[o] - alpha * AP[o];
    R[o] = ri;
    const double zi = diag ? ri / diag[i] : ri;
    acc[0] += ri * ri;
    acc[1] += ri * zi;
  }
  store_rhs_partials<2>(acc, ln, K, part);
}

// Loop test (cg.py:126-129) and the scalars of Update per right-hand side.
__global__ __launch_bounds__(kRB) void k_m_test(int K, int64_t g, const double *__restrict__ part,
                                               MrhsState *st, int32_t *active) {
  const int k = blockIdx.x;
  MrhsState *s = st + k;
  if (s->done) return;
  double rr[1], rz[1];
  sum_partials<1>(part, 0, (int64_t)k * g, g, rr);
  sum_partials<1>(part, 0, ((int64_t)K + k) * g, g, rz);
  if (threadIdx.x != 0) return;
  const double rnorm = cr_sqrt(rr[0]);
  const int64_t kk = s->k_cur;
  s->rr = rr[0];
  s->rz_new = rz[0];
  s->rnorm = rnorm;
  const bool stop = ((rnorm == rnorm) && rnorm <= s->tol) ||
                    (s->maxiter >= 0 && kk >= s->maxiter) || (rnorm != rnorm && s->maxiter < 0);
  if (stop) {
    s->done = 1;
    s->iters = kk;
    atomicSub(active, 1);
    return;
  }
  s->beta = rz[0] / s->rz_cur;
  s->k_next = kk + 1;
  s->rz_next = rz[0];
}

// (Two rows per loop iteration in this kernel and k_m_xr, both rows' loads
// in flight together, measured 2 % slower: 5.21 vs 5.11 s for 64 right-hand
// sides on 1024^2.)
// The step's x += alpha p (cg.py:46, for every side that ran this step,
// including one the loop test has just stopped) and Update (cg.py:67-71):
// p = z + beta p with z = M r, for the sides still running.
__global__ __launch_bounds__(kRB) void k_m_p(int64_t n, int K, const double *__restrict__ diag,
                                            double *__restrict__ X, const double *__restrict__ R,
                                            double *__restrict__ P, const MrhsState *st) {
  const Lanes ln = lanes_of(K);
  const MrhsState *s = st + (ln.k < K ? ln.k : 0);
  const bool xon = ln.k < K && s->xpend;
  const bool pon = xon && !s->done;
  const double alpha = s->alpha, beta = s->beta;
  ESR_MRHS_ROWS(i, n, ln) {
    if (!xon) continue;
    const int64_t o = i * K + ln.k;
    const double pi = P[o];
    X[o] = X[o] + alpha * pi;
    if (pon) {
      const double ri = R[o];
      const double zi = diag ? ri / diag[i] : ri;
      P[o] = zi + beta * pi;
    }
  }
}

inline unsigned mrhs_grid_cap(int64_t n, int K, int64_t cap) {
  int L = 1;
  while (L < K) L <<= 1;
  const int64_t g = ceil_div(n, kRB / L);
  return (unsigned)(g < cap ? (g > 0 ? g : 1) : cap);
}
inline unsigned mrhs_grid(int64_t n, int K) { return mrhs_grid_cap(n, K, kMaxGridM); }
// The grid of the step kernels: the product's wavefront serves one row at a
// time through dependent row-pointer -> column -> neighbour-row loads, and
// the update kernels stream 3-5 vectors of n K values, so all of them need
// many wavefronts in flight; their partials (per workgroup and right-hand
// side) are summed by k_m_alpha / k_m_test in a fixed order whatever their
// number. (The update kernels on the 512-workgroup grid of k_m_init: 64
// right-hand sides on 1024^2 in 6.03 s; on this grid 5.47 s.)
inline unsigned mrhs_spmv_grid(int64_t n, int K) { return mrhs_grid_cap(n, K, kMaxGridSpmvM); }

} // namespace
} // namespace esr

using namespace esr;

ESR_API int esr_mrhs_sizes(int64_t n, int64_t K, int64_t *state_bytes, int64_t *ws_bytes) {
  ESR_REQUIRE(n >= 1 && K >= 1 && K <= kMaxK && state_bytes && ws_bytes,
              "esr_mrhs_sizes: need n >= 1 and 1 <= K <= %d", kMaxK);
  *state_bytes = K * (int64_t)sizeof(MrhsState);
  const int64_t g = mrhs_grid(n, (int)K), g1 = mrhs_spmv_grid(n, (int)K);
  *ws_bytes = K * (2 * g > 2 * g1 ? 2 * g : 2 * g1) * (int64_t)sizeof(double);
  return ESR_OK;
}

// B, X, R, Z, P, AP: (n, K) row-major float64; state: K MrhsState with rtol,
// atol, maxiter set by the caller; active: one int32 (set to K).
ESR_API int esr_mrhs_init(int64_t n, int64_t K, const int32_t *rowptr, const int32_t *col,
                          const double *Ae, const double *Av, const double *diag,
                          const double *B, const double *X, double *R, double *Z, double *P,
                          void *state, double *ws, int32_t *active, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1 && K >= 1 && K <= kMaxK, "esr_mrhs_init: bad sizes");
  ESR_REQUIRE(rowptr && col && Ae && Av && B && X && R && Z && P && state && ws && active,
              "esr_mrhs_init: null argument");
  hipStream_t s = as_stream(stream);
  const unsigned g = mrhs_grid(n, (int)K);
  k_m_init<<<g, kRB, 0, s>>>(n, (int)K, rowptr, col, Ae, Av, diag, B, X, R, Z, P, ws);
  k_m_init_fin<<<(unsigned)K, kRB, 0, s>>>((int)K, g, ws, (MrhsState *)state, active);
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

namespace {
// The step after the product: alpha, r, loop test, x and p.
void mrhs_step_rest(int64_t n, int K, const double *diag, double *X, double *R, double *P,
                    const double *AP, MrhsState *st, double *ws, int32_t *active, unsigned g1,
                    hipStream_t s) {
  k_m_alpha<<<(unsigned)K, kRB, 0, s>>>(K, g1, ws, st);
  k_m_xr<<<g1, kRB, 0, s>>>(n, K, diag, R, AP, st, ws);
  k_m_test<<<(unsigned)K, kRB, 0, s>>>(K, g1, ws, st, active);
  k_m_p<<<g1, kRB, 0, s>>>(n, K, diag, X, R, P, st);
}
} // namespace

// `count` iterations; each is a no-op for a converged right-hand side.
ESR_API int esr_mrhs_steps(int64_t n, int64_t K, const int32_t *rowptr, const int32_t *col,
                           const double *Ae, const double *Av, const double *diag, double *X,
                           double *R, double *Z, double *P, double *AP, void *state, double *ws,
                           int32_t *active, int64_t count, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1 && K >= 1 && K <= kMaxK && count >= 0, "esr_mrhs_steps: bad sizes");
  ESR_REQUIRE(rowptr && col && Ae && Av && X && R && Z && P && AP && state && ws && active,
              "esr_mrhs_steps: null argument");
  hipStream_t s = as_stream(stream);
  const unsigned g = mrhs_grid(n, (int)K), g1 = mrhs_spmv_grid(n, (int)K);
  MrhsState *st = (MrhsState *)state;
  for (int64_t it = 0; it < count; ++it) {
    k_m_spmv<<<g1, kRB, 0, s>>>(n, (int)K, rowptr, col, Ae, Av, P, AP, st, ws);
    mrhs_step_rest(n, (int)K, diag, X, R, P, AP, st, ws, active, g1, s);
    ESR_LAUNCHED();
  }
  return ESR_OK;
  ESR_GUARD_END
}

// esr_mrhs_steps on the padded ELL copy of the operator (rows of at most
// ESR_MRHS_ELL_W entries; esr_mrhs_ell builds it): the same results.
ESR_API int esr_mrhs_steps_ell(int64_t n, int64_t K, const uint8_t *len, const int32_t *ecol,
                               const double *eval, const double *Av, const double *diag,
                               double *X, double *R, double *Z, double *P, double *AP,
                               void *state, double *ws, int32_t *active, int64_t count,
                               void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1 && K >= 1 && K <= kMaxK && count >= 0, "esr_mrhs_steps_ell: bad sizes");
  ESR_REQUIRE(len && ecol && eval && Av && X && R && Z && P && AP && state && ws && active,
              "esr_mrhs_steps_ell: null argument");
  hipStream_t s = as_stream(stream);
  const unsigned g1 = mrhs_spmv_grid(n, (int)K);
  MrhsState *st = (MrhsState *)state;
  for (int64_t it = 0; it < count; ++it) {
    k_m_spmv_ell<<<g1, kRB, 0, s>>>(n, (int)K, len, ecol, eval, Av, P, AP, st, ws);
    mrhs_step_rest(n, (int)K, diag, X, R, P, AP, st, ws, active, g1, s);
    ESR_LAUNCHED();
  }
  return ESR_OK;
  ESR_GUARD_END
}

namespace {
__global__ __launch_bounds__(kRB) void k_m_ell(int64_t n, const int32_t *__restrict__ rowptr,
                                              const int32_t *__restrict__ col,
                                              const double *__restrict__ Ae,
                                              uint8_t *__restrict__ len,
                                              int32_t *__restrict__ ecol,
                                              double *__restrict__ eval) {
  const int64_t i = (int64_t)blockIdx.x * kRB + threadIdx.x;
  if (i >= n) return;
  const int32_t b = rowptr[i], m = rowptr[i + 1] - b;
  len[i] = (uint8_t)m;
#pragma unroll
  for (int k = 0; k < kEllW; ++k) {
    ecol[i * kEllW + k] = k < m ? col[b + k] : (int32_t)i;
    eval[i * kEllW + k] = k < m ? Ae[b + k] : 0.0;
  }
}
} // namespace

// Padded ELL copy of a CSR operator whose rows hold at most ESR_MRHS_ELL_W
// entries (the caller checks the row lengths: a longer row would be cut).
ESR_API int esr_mrhs_ell(int64_t n, const int32_t *rowptr, const int32_t *col, const double *Ae,
                         uint8_t *len, int32_t *ecol, double *eval, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1 && rowptr && col && Ae && len && ecol && eval,
              "esr_mrhs_ell: bad arguments");
  k_m_ell<<<(unsigned)ceil_div(n, kRB), kRB, 0, as_stream(stream)>>>(n, rowptr, col, Ae, len,
                                                                     ecol, eval);
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}
