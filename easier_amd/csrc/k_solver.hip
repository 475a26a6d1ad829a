// Fused Krylov iterations over a CSR Linsys operator.
//
// The operator is Linsys (easier/numeric/linsys.py:23-24):
//   y[r] = (sum_{j in row r, in edge order} x[col[j]] * Ae[j]) + x[r] * Av[r]
// with the row sum folded from 0 left to right, which is what the reference's
// Reducer computes for the renumbered, row-sorted edge list.
//
// Dot products are accumulated in double, one partial per workgroup, and
// summed in a fixed order by every workgroup of the kernel that consumes
// them (esr_krylov.h): no atomics, no fences, the same bits on every run.
// Solver scalars stay on the device; every kernel of an iteration is a
// no-op once the loop has exited, so chunks of iterations can be captured
// in a HIP graph and replayed until the host sees `done`.
#include <algorithm>
#include <type_traits>
#include <utility>

#include "esr_common.h"
#include "esr_krylov.h"
#include "esr_solver.h"

namespace esr {
namespace {

// ---------------------------------------------------------------------------
// PCG
// ---------------------------------------------------------------------------
// Init (cg.py:25-29): r = b - A x; z = M r; p = z; partials of r.z and b.b.
template <typename T>
__global__ __launch_bounds__(kRB) void k_pcg_init(
    int64_t n, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const T *__restrict__ Ae, const T *__restrict__ Av, const T *__restrict__ diag,
    const T *__restrict__ bvec, const T *__restrict__ x, T *__restrict__ r,
    T *__restrict__ z, T *__restrict__ p, double *__restrict__ part) {
  double acc[2] = {0.0, 0.0};
  ESR_GRID_ROWS(i, n) {
    const T ax = linsys_row<T>(i, rowptr, col, Ae, Av, [&](int64_t c) { return x[c]; });
    const T ri = bvec[i] - ax;
    const T zi = diag ? ri / diag[i] : ri;
    r[i] = ri;
    z[i] = zi;
    p[i] = zi;
    acc[0] += (double)(ri * zi);
    acc[1] += (double)(bvec[i] * bvec[i]);
  }
  store_partials<2>(acc, part);
}

// bnorm = ||b||; tol = max(rtol*bnorm, atol) if atol else rtol*bnorm
// (cg.py:112-115); loop counters reset; p (in buffer 0) is the first
// direction as it is. One workgroup.
__global__ __launch_bounds__(kRB) void k_pcg_init_fin(int64_t g, const double *__restrict__ part,
                                                     PcgState *st) {
  double tot[2];
  sum_partials<2>(part, g, 0, g, tot);
  if (threadIdx.x == 0) {
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
    st->fresh = 1;
    st->pbuf = 0;
    st->pbuf_next = 0;
  }
}

// Loop test (cg.py:126-129) of the step just completed, from the partials
// of r.r and r.z its second kernel stored: rnorm = ||r|| <= tol, or the
// step cap, or a NaN residual without a cap. On exit the leading thread
// records the exit (the same values whichever kernel finds it first).
struct PcgTest {
  double rr, rz, rnorm;
  int64_t k;
  bool exit;
};
__device__ __forceinline__ PcgTest pcg_test(const double *__restrict__ part, int64_t g,
                                            int64_t k, PcgState *st, bool lead) {
  double s[2];
  sum_partials<2>(part, g, 0, g, s);
  PcgTest t;
  t.rr = s[0];
  t.rz = s[1];
  t.rnorm = cr_sqrt(s[0]);
  t.k = k;
  const bool converged = (t.rnorm == t.rnorm) && t.rnorm <= st->tol;
  const bool capped = st->maxiter >= 0 && t.k >= st->maxiter;
  const bool nan_stop = t.rnorm != t.rnorm && st->maxiter < 0;
  t.exit = converged || capped || nan_stop;
  if (t.exit && lead) {
    st->rr = t.rr;
    st->rz_new = t.rz;
    st->rnorm = t.rnorm;
    st->done = 1;
    st->iters = t.k;
  }
  return t;
}

// One PCG iteration is two kernels on small systems, three on large ones
// (kFuseUpdateMax); both forms run the same operations on the same values.
//
// At 4.2 M vertices the three-kernel iteration was 170 us: kernel 1 103 us
// (SpMV, 4.1 TB/s of its 420 MB), kernel 2 48 us (5.6 TB/s), the Update
// 19 us (5.2 TB/s); with the wave-uniform slot count of the product it is
// 158 us. Measured alternatives for the SpMV, all slower or equal:
// groups of 16 or 64 row tiles per XCD (1.539-1.543 s per solve, unchanged),
// a contiguous chunk of rows per workgroup (1.62 s), each wavefront's CSR
// segment staged in LDS before the gathers (1.56 s; float32 1.49 vs 1.31 s).
//
// Kernel 1: the loop test and Update (cg.py:67-71, 126-129) of the previous
// step -- beta = rz_new / rz and p = z + beta p -- fused with Step part 1
// (cg.py:44-47), Ap = A p and partials of p.Ap. The new direction is
// evaluated where it is used, z[c] + beta p[c] for the row and every
// neighbour (the same operations, so the same bits, as updating p first),
// read from one direction buffer and written to the other, so no workgroup
// overwrites a value another one still gathers. After Init, a re-arm or the
// separate Update kernel (fresh) the step starts from p as it is.
template <typename T, bool FUSED>
__global__ __launch_bounds__(kRB) void k_pcg_step1(
    int64_t n, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const T *__restrict__ Ae, const T *__restrict__ Av, const T *__restrict__ z,
    T *__restrict__ p0, T *__restrict__ p1, T *__restrict__ Ap, PcgState *st,
    const double *__restrict__ part_in, double *__restrict__ part_out) {
  if (st->done) return;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const int64_t cur = st->pbuf;
  const T *__restrict__ pin = cur ? p1 : p0;
  T *__restrict__ pout = cur ? p0 : p1;
  double acc[1] = {0.0};
  if (!FUSED || st->fresh) {
    if (lead) {
      st->k_next = st->k_cur;
      st->rz_next = st->rz_cur;
      st->pbuf_next = cur;
    }
    ESR_GRID_ROWS(i, n) {
      const T y = linsys_row_wave<T>(i, rowptr, col, Ae, Av, [&](int64_t c) { return pin[c]; });
      Ap[i] = y;
      acc[0] += (double)(pin[i] * y);
    }
  } else if constexpr (FUSED) {
    const PcgTest t = pcg_test(part_in, gridDim.x, st->k_cur, st, lead);
    if (t.exit) return;
    const T beta = (T)(t.rz / st->rz_cur);
    if (lead) {
      st->rr = t.rr;
      st->rz_new = t.rz;
      st->rnorm = t.rnorm;
      st->k_next = t.k + 1;
      st->rz_next = t.rz;
      st->beta = (double)beta;
      st->pbuf_next = 1 - cur;
    }
    ESR_GRID_ROWS(i, n) {
      const T y = linsys_row<T>(i, rowptr, col, Ae, Av,
                                [&](int64_t c) { return z[c] + beta * pin[c]; });
      const T pi = z[i] + beta * pin[i];
      pout[i] = pi;
      Ap[i] = y;
      acc[0] += (double)(pi * y);
    }
  }
  store_partials<1>(acc, part_out);
}

// Kernel 1 on the operator's SELL-64 form (esr_krylov.h): the same
// decisions, values and fold as k_pcg_step1 (rows are visited 64 at a time,
// so a wavefront's rows are one slice), bit for bit the same results.
template <typename T, typename C, bool FUSED>
__global__ __launch_bounds__(kRB) void k_pcg_step1_sell(
    int64_t n, const int64_t *__restrict__ soff, const uint8_t *__restrict__ cnt,
    const C *__restrict__ colv, const T *__restrict__ val, const T *__restrict__ Av,
    const T *__restrict__ z, T *__restrict__ p0, T *__restrict__ p1, T *__restrict__ Ap,
    PcgState *st, const double *__restrict__ part_in, double *__restrict__ part_out) {
  if (st->done) return;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const int64_t cur = st->pbuf;
  const T *__restrict__ pin = cur ? p1 : p0;
  T *__restrict__ pout = cur ? p0 : p1;
  double acc[1] = {0.0};
  if (!FUSED || st->fresh) {
    if (lead) {
      st->k_next = st->k_cur;
      st->rz_next = st->rz_cur;
      st->pbuf_next = cur;
    }
    ESR_GRID_ROWS(i, n) {
      const T y = sell_row<T>(i, soff, cnt, colv, val, Av, [&](int64_t c) { return pin[c]; });
      Ap[i] = y;
      acc[0] += (double)(pin[i] * y);
    }
  } else if constexpr (FUSED) {
    const PcgTest t = pcg_test(part_in, gridDim.x, st->k_cur, st, lead);
    if (t.exit) return;
    const T beta = (T)(t.rz / st->rz_cur);
    if (lead) {
      st->rr = t.rr;
      st->rz_new = t.rz;
      st->rnorm = t.rnorm;
      st->k_next = t.k + 1;
      st->rz_next = t.rz;
      st->beta = (double)beta;
      st->pbuf_next = 1 - cur;
    }
    ESR_GRID_ROWS(i, n) {
      const T y = sell_row<T>(i, soff, cnt, colv, val, Av,
                              [&](int64_t c) { return z[c] + beta * pin[c]; });
      const T pi = z[i] + beta * pin[i];
      pout[i] = pi;
      Ap[i] = y;
      acc[0] += (double)(pi * y);
    }
  }
  store_partials<1>(acc, part_out);
}

// Kernel 2: Step part 2 (cg.py:47-52) fused with the first half of Update
// (cg.py:67): alpha = rz / pAp; x += alpha p; r -= alpha Ap; z = M r;
// partials of r.r, r.z. Workgroup 0 commits what kernel 1 decided.
template <typename T>
__global__ __launch_bounds__(kRB) void k_pcg_step2(int64_t n, const T *__restrict__ diag,
                                                  T *__restrict__ x, T *__restrict__ r,
                                                  T *__restrict__ z, const T *__restrict__ p0,
                                                  const T *__restrict__ p1,
                                                  const T *__restrict__ Ap, PcgState *st,
                                                  const double *__restrict__ part_in,
                                                  int64_t g_in, double *__restrict__ part_out) {
  if (st->done) return;
  double pAp[1];
  sum_partials<1>(part_in, g_in, 0, g_in, pAp);
  const double rz = st->rz_next;
  const double alpha_d = rz / pAp[0];
  const T *__restrict__ p = st->pbuf_next ? p1 : p0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->pAp = pAp[0];
    st->alpha = alpha_d;
    st->k_cur = st->k_next;
    st->rz_cur = rz;
    st->pbuf = st->pbuf_next;
    st->fresh = 0;
  }
  const T alpha = (T)alpha_d;
  double acc[2] = {0.0, 0.0};
  ESR_GRID_ROWS(i, n) {
    const T pi = p[i];
    const T ri = r[i] - alpha * Ap[i];
    x[i] = x[i] + alpha * pi;
    r[i] = ri;
    const T zi = diag ? ri / diag[i] : ri;
    z[i] = zi;
    acc[0] += (double)(ri * ri);
    acc[1] += (double)(ri * zi);
  }
  store_partials<2>(acc, part_out);
}

// Kernel 3 of the three-kernel form: the loop test and Update of the step
// just completed, p = z + beta p in place. It reads the *_next copies (equal
// to *_cur after kernel 2) and writes *_cur, so the next kernel 1 starts
// fresh from the updated p.
template <typename T>
__global__ __launch_bounds__(kRB) void k_pcg_update(int64_t n, const T *__restrict__ z,
                                                   T *__restrict__ p0, T *__restrict__ p1,
                                                   PcgState *st,
                                                   const double *__restrict__ part_in) {
  if (st->done) return;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const PcgTest t = pcg_test(part_in, gridDim.x, st->k_next, st, lead);
  if (t.exit) return;
  const T beta = (T)(t.rz / st->rz_next);
  T *__restrict__ p = st->pbuf_next ? p1 : p0;
  ESR_GRID_ROWS(i, n) p[i] = z[i] + beta * p[i];
  if (lead) {
    st->rr = t.rr;
    st->rz_new = t.rz;
    st->rnorm = t.rnorm;
    st->beta = (double)beta;
    st->k_cur = t.k + 1;
    st->rz_cur = t.rz;
    st->fresh = 1;
  }
}

// End of a batch in the two-kernel form: the loop test of the last step, so
// that the host sees `done` without another batch (kernel 1 of a following
// step re-derives the same exit from the same partials), and once the loop
// has exited the last direction is copied into the caller's p if it lives
// in the second buffer (the reference's p after the loop; no workgroup
// changes pbuf, so all decide alike).
template <typename T>
__global__ __launch_bounds__(kRB) void k_pcg_settle(int64_t n, T *__restrict__ p0,
                                                   const T *__restrict__ p1,
                                                   PcgState *st,
                                                   const double *__restrict__ part) {
  if (st->fresh) return;
  const bool exited =
      st->done || pcg_test(part, gridDim.x, st->k_cur, st, blockIdx.x == 0 && threadIdx.x == 0).exit;
  if (exited && st->pbuf == 1) ESR_GRID_ROWS(i, n) p0[i] = p1[i];
}

// Re-arm after a reliable update (the caller replaced r by the true residual
// of the current iterate): z = M r and partials of r.z. k_pcg_rearm_fin
// takes the Update of cg.py:67-71 with this residual -- beta = r.z / rz of
// the last step -- and installs the new exit tolerance; k_pcg_rearm_p sets
// p = z + beta p in the buffer that holds p, so the search directions stay
// conjugate, and the next step starts from it as it is.
template <typename T>
__global__ __launch_bounds__(kRB) void k_pcg_rearm(int64_t n, const T *__restrict__ diag,
                                                  const T *__restrict__ r, T *__restrict__ z,
                                                  double *__restrict__ part) {
  double acc[1] = {0.0};
  ESR_GRID_ROWS(i, n) {
    const T ri = r[i];
    const T zi = diag ? ri / diag[i] : ri;
    z[i] = zi;
    acc[0] += (double)(ri * zi);
  }
  store_partials<1>(acc, part);
}

__global__ __launch_bounds__(kRB) void k_pcg_rearm_fin(int64_t g, const double *__restrict__ part,
                                                      PcgState *st, double tol) {
  double tot[1];
  sum_partials<1>(part, g, 0, g, tot);
  if (threadIdx.x == 0) {
    st->beta = tot[0] / st->rz_cur;
    st->rz_new = tot[0];
    st->rz_cur = tot[0];
    st->rz_next = tot[0];
    st->tol = tol;
    st->k_cur = st->k_cur + 1;
    st->k_next = st->k_cur;
    st->fresh = 1;
    st->pbuf_next = st->pbuf;
    st->done = 0;
  }
}

template <typename T>
__global__ __launch_bounds__(kRB) void k_pcg_rearm_p(int64_t n, const T *__restrict__ z,
                                                    T *__restrict__ p0, T *__restrict__ p1,
                                                    const PcgState *st) {
  const T beta = (T)st->beta;
  T *__restrict__ p = st->pbuf ? p1 : p0;
  ESR_GRID_ROWS(i, n) p[i] = z[i] + beta * p[i];
}

// ---------------------------------------------------------------------------
// GMRES Arnoldi cycle
// ---------------------------------------------------------------------------
// Partials of ||b||^2 (Init, gmres.py:36-37); k_gm_bnorm_fin finishes.
template <typename T>
__global__ __launch_bounds__(kRB) void k_gm_bnorm(int64_t n, const T *__restrict__ bvec,
                                                 double *__restrict__ part) {
  double acc[1] = {0.0};
  ESR_GRID_ROWS(i, n) acc[0] += (double)(bvec[i] * bvec[i]);
  store_partials<1>(acc, part);
}

__global__ __launch_bounds__(kRB) void k_gm_bnorm_fin(int64_t g, const double *__restrict__ part,
                                                     GmresState *st) {
  double tot[1];
  sum_partials<1>(part, g, 0, g, tot);
  if (threadIdx.x == 0) st->bnorm = cr_sqrt(tot[0]);
}

// UpdateRnorm (gmres.py:22-25): r = b - A x; partials of r.r;
// k_gm_residual_fin finishes ||r|| and clears the cycle's breakdown flag.
template <typename T>
__global__ __launch_bounds__(kRB) void k_gm_residual(
    int64_t n, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const T *__restrict__ Ae, const T *__restrict__ Av, const T *__restrict__ bvec,
    const T *__restrict__ x, T *__restrict__ r, double *__restrict__ part,
    const GmresState *loop) {
  if (loop && loop->done) return;
  double acc[1] = {0.0};
  ESR_GRID_ROWS(i, n) {
    const T ri = bvec[i] - linsys_row<T>(i, rowptr, col, Ae, Av, [&](int64_t c) { return x[c]; });
    r[i] = ri;
    acc[0] += (double)(ri * ri);
  }
  store_partials<1>(acc, part);
}

__global__ __launch_bounds__(kRB) void k_gm_residual_fin(int64_t g,
                                                        const double *__restrict__ part,
                                                        GmresState *st) {
  double tot[1];
  sum_partials<1>(part, g, 0, g, tot);
  if (threadIdx.x == 0) {
    st->rnorm = cr_sqrt(tot[0]);
    st->broken = 0;
    st->jend = -1;
  }
}

// InitV/UpdateV + InitW (gmres.py:48, 64, 120): V_j = src / s and
// w = A(M(V_j)), M = identity or r / diag. For j = 0, s = ||r||. For j > 0,
// s = H[j, j-1] = ||w|| from the previous norm kernel's partials (stored in
// H, gmres.py:90, 210-213; g_in of them); below 1e-15 the cycle breaks down at j - 1. The
// breakdown flag is read and (by workgroup 0) written in this kernel, so both
// accesses are atomic; every workgroup derives the same decision either way.
template <typename T>
__global__ __launch_bounds__(kRB) void k_gm_spmv(
    int64_t n, int j, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const T *__restrict__ Ae, const T *__restrict__ Av, const T *__restrict__ diag,
    const T *__restrict__ src, T *__restrict__ Vj, T *__restrict__ w, GmresState *st,
    const double *__restrict__ part_in, int64_t g_in) {
  if (__hip_atomic_load(&st->broken, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  T s;
  if (j == 0) {
    s = (T)st->rnorm;
  } else {
    double tot[1];
    sum_partials<1>(part_in, g_in, 0, g_in, tot);
    const double hn = (double)(T)cr_sqrt(tot[0]);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->H[j * kGmresMaxRestart + (j - 1)] = hn;
      st->jend = j - 1;
      if (hn < 1e-15)
        __hip_atomic_store(&st->broken, (int64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (hn < 1e-15) return;
    s = (T)hn;
  }
  auto mv = [&](int64_t c) {
    const T v = src[c] / s;
    return diag ? v / diag[c] : v;
  };
  ESR_GRID_ROWS(i, n) {
    Vj[i] = src[i] / s;
    w[i] = linsys_row<T>(i, rowptr, col, Ae, Av, mv);
  }
}

// Orthogonalisation of w against V_0..V_j (gmres.py:79-90 computes
// H[i, j] = V_i . w and w -= H[i, j] V_i one i at a time, modified
// Gram-Schmidt). Here: classical Gram-Schmidt applied twice (CGS2), which
// keeps the basis as orthogonal as MGS does (Giraud, Langou, Rozloznik and
// van den Eshof, Numer. Math. 101, 2005) with three grid-wide reductions per
// Arnoldi step instead of j + 2, each one kernel:
//   k_gm_dots:     partials of h_i = V_i . w, i <= j
//   k_gm_cgs<1>:   h_i from them, w -= sum_i h_i V_i, partials of g_i = V_i . w
//   k_gm_cgs<0>:   g_i, w -= sum_i g_i V_i, H[i, j] = h_i + g_i, partials of w . w
// (the next step's k_gm_spmv finishes ||w|| = H[j+1, j]). In exact
// arithmetic both give the same H and basis. The kernels run on a grid of at
// most kGsMaxGrid workgroups so that re-summing (j + 1) rows of partials in
// every workgroup stays cheap. JB is the register block of basis vectors per
// row (8, 16 or 32 >= j + 1).
constexpr int kGsMaxGrid = 256;
inline unsigned gs_grid(int64_t n) {
  const int64_t g = ceil_div(n, kRB);
  return (unsigned)(g < kGsMaxGrid ? (g > 0 ? g : 1) : kGsMaxGrid);
}

// Workgroup sums of per-thread values acc[i] (i < cnt), stored to
// part[i * gridDim.x + blockIdx.x]. Within a wavefront the JB sums are
// formed together by halving exchanges: at offset 32 a lane keeps one half
// of its values and adds its partner's copy of that half (lanes with bit 5
// set keep the upper half), then offset 16 on the remaining half, ... until
// one value is left, which is value lane / (64 / JB) summed over 64 / (64 /
// JB) lanes; xor exchanges over the remaining low offsets finish it. JB - 1
// + log2(64 / JB) exchanges instead of 6 per value; a fixed order, so every
// run gives the same bits.
template <int JB>
__device__ __forceinline__ void gs_store_partials(double (&acc)[JB], int cnt,
                                                  double *__restrict__ part) {
  static_assert(JB >= 1 && JB <= 32 && (JB & (JB - 1)) == 0, "JB: a power of 2 up to 32");
  __shared__ double sh[JB][kRB / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  constexpr int kLog = JB >= 32 ? 5 : JB >= 16 ? 4 : JB >= 8 ? 3 : JB >= 4 ? 2 : JB >= 2 ? 1 : 0;
#pragma unroll
  for (int step = 0; step < kLog; ++step) {
    const int m = JB >> (step + 1), off = 32 >> step;
    const bool upper = (lane & off) != 0;
#pragma unroll
    for (int k = 0; k < m; ++k) {
      // select between values, not between array slots: a select of the
      // two slots' addresses would put acc in scratch memory
      const double lo = acc[k], hi = acc[k + m];
      const double send = upper ? lo : hi;
      const double keep = upper ? hi : lo;
      acc[k] = keep + __shfl_xor(send, off, 64);
    }
  }
#pragma unroll
  for (int off = 32 >> kLog; off >= 1; off >>= 1) acc[0] += __shfl_xor(acc[0], off, 64);
  constexpr int span = 64 / JB;  // lanes that hold the same sum
  if (lane % span == 0) sh[lane / span][wid] = acc[0];
  __syncthreads();
  if ((int)threadIdx.x < cnt) {
    double t = 0.0;
    for (int w = 0; w < kRB / 64; ++w) t += sh[threadIdx.x][w];
    part[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = t;
  }
}

// out[i] = (T) sum_{b < g} part[i * g + b] for i < cnt (g <= kGsMaxGrid),
// in every workgroup: wavefront w sums rows i = w, w + 4, ...; each lane
// adds partials b = lane, lane + 64, ... in that order, then the shuffle
// tree (a fixed order, the same in every workgroup). All loads (clamped
// positions) are issued before the first sum.
template <typename T, int JB>
__device__ __forceinline__ void gs_sum_rows(const double *__restrict__ part, int64_t g, int cnt,
                                            T *out) {
  constexpr int RW = (JB + kRB / 64 - 1) / (kRB / 64);  // rows per wavefront
  constexpr int KB = kGsMaxGrid / 64;                   // partials per lane and row
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double v[RW][KB];
#pragma unroll
  for (int q = 0; q < RW; ++q) {
    const int i = wid + (kRB / 64) * q;
    const int64_t ic = i < cnt ? i : cnt - 1;
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      const int64_t b = lane + 64 * k;
      v[q][k] = part[ic * g + (b < g ? b : g - 1)];
    }
  }
#pragma unroll
  for (int q = 0; q < RW; ++q) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < KB; ++k)
      if (lane + 64 * k < g) t += v[q][k];
    t = kr_wave_sum(t);
    const int i = wid + (kRB / 64) * q;
    if (lane == 0 && i < cnt) out[i] = (T)t;
  }
  __syncthreads();
}

template <typename T, int JB>
__global__ __launch_bounds__(kRB) void k_gm_dots(int64_t n, int j, const T *__restrict__ V,
                                                const T *__restrict__ w, const GmresState *st,
                                                double *__restrict__ part_out) {
  if (st->broken) return;
  double acc[JB];
#pragma unroll
  for (int i = 0; i < JB; ++i) acc[i] = 0.0;
  ESR_GRID_ROWS(r, n) {
    const T wr = w[r];
    // every load unconditional (rows past j clamped to row j), so all of
    // them are in flight together; the sums past j are never stored
#pragma unroll
    for (int i = 0; i < JB; ++i)
      acc[i] += (double)(V[(int64_t)(i <= j ? i : j) * n + r] * wr);
  }
  gs_store_partials<JB>(acc, j + 1, part_out);
}

// k_gm_spmv and k_gm_dots as one kernel (systems of at most kFuseUpdateMax
// rows, where the Arnoldi step is launch-bound): on the dots kernel's grid
// and rows, each thread forms V_j[r] and w[r] = A(M(V_j))[r], then adds
// V_i[r] w[r] (i < j from memory, i >= j its own V_j[r]) to the partials of
// h_i -- the values, products and summation order of the two kernels, so
// the same bits.
template <typename T, int JB>
__global__ __launch_bounds__(kRB) void k_gm_spmv_dots(
    int64_t n, int j, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const T *__restrict__ Ae, const T *__restrict__ Av, const T *__restrict__ diag,
    const T *__restrict__ src, T *__restrict__ V, T *__restrict__ w, GmresState *st,
    const double *__restrict__ part_in, int64_t g_in, double *__restrict__ part_out) {
  if (__hip_atomic_load(&st->broken, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  T s;
  if (j == 0) {
    s = (T)st->rnorm;
  } else {
    double tot[1];
    sum_partials<1>(part_in, g_in, 0, g_in, tot);
    const double hn = (double)(T)cr_sqrt(tot[0]);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->H[j * kGmresMaxRestart + (j - 1)] = hn;
      st->jend = j - 1;
      if (hn < 1e-15)
        __hip_atomic_store(&st->broken, (int64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (hn < 1e-15) return;
    s = (T)hn;
  }
  auto mv = [&](int64_t c) {
    const T v = src[c] / s;
    return diag ? v / diag[c] : v;
  };
  T *__restrict__ Vj = V + (int64_t)j * n;
  double acc[JB];
#pragma unroll
  for (int i = 0; i < JB; ++i) acc[i] = 0.0;
  ESR_GRID_ROWS(r, n) {
    const T vj = src[r] / s;
    Vj[r] = vj;
    const T wr = linsys_row<T>(r, rowptr, col, Ae, Av, mv);
    w[r] = wr;
#pragma unroll
    for (int i = 0; i < JB; ++i) {
      const T vi = i < j ? V[(int64_t)i * n + r] : vj;
      acc[i] += (double)(vi * wr);
    }
  }
  gs_store_partials<JB>(acc, j + 1, part_out);
}

template <typename T, int JB, bool FIRST>
__global__ __launch_bounds__(kRB) void k_gm_cgs(int64_t n, int j, const T *__restrict__ V,
                                               T *__restrict__ w, GmresState *st,
                                               const double *__restrict__ part_in,
                                               double *__restrict__ part_out) {
  if (st->broken) return;
  __shared__ T h[JB];
  gs_sum_rows<T, JB>(part_in, gridDim.x, j + 1, h);
  if (blockIdx.x == 0 && (int)threadIdx.x <= j) {
    double *Hij = &st->H[threadIdx.x * kGmresMaxRestart + j];
    // pass 1 stores h_i; pass 2 completes H[i, j] = h_i + g_i
    *Hij = FIRST ? (double)h[threadIdx.x] : (double)((T)*Hij + h[threadIdx.x]);
  }
  double acc[FIRST ? JB : 1];
#pragma unroll
  for (int i = 0; i < (FIRST ? JB : 1); ++i) acc[i] = 0.0;
  ESR_GRID_ROWS(r, n) {
    T v[JB];
#pragma unroll
    for (int i = 0; i < JB; ++i) v[i] = V[(int64_t)(i <= j ? i : j) * n + r];
    T wr = w[r];
#pragma unroll
    for (int i = 0; i < JB; ++i)
      if (i <= j) wr = wr - h[i] * v[i];
    w[r] = wr;
    if (FIRST) {
#pragma unroll
      for (int i = 0; i < JB; ++i)
        if (i <= j) acc[i] += (double)(v[i] * wr);
    } else {
      acc[0] += (double)(wr * wr);
    }
  }
  gs_store_partials<FIRST ? JB : 1>(acc, FIRST ? j + 1 : 1, part_out);
}

// End of a cycle of m steps: H[m, m-1] = ||w|| from the last norm's partials.
template <typename T>
__global__ __launch_bounds__(kRB) void k_gm_fin(int64_t g, int m, GmresState *st,
                                               const double *__restrict__ part_in) {
  if (st->broken) return;
  double tot[1];
  sum_partials<1>(part_in, g, 0, g, tot);
  if (threadIdx.x == 0) {
    const double hn = (double)(T)cr_sqrt(tot[0]);
    st->H[m * kGmresMaxRestart + (m - 1)] = hn;
    st->jend = m - 1;
    if (hn < 1e-15) st->broken = 1;
  }
}

// UpdateX (gmres.py:133-136): x += M(sum_{i<jc} V_i y_i).
template <typename T>
__global__ __launch_bounds__(kRB) void k_gm_update_x(int64_t n, int jc, const T *__restrict__ V,
                                                    const T *__restrict__ y,
                                                    const T *__restrict__ diag,
                                                    T *__restrict__ x) {
  ESR_GRID_ROWS(r, n) {
    T dx = T(0);
    for (int i = 0; i < jc; ++i) dx = dx + V[(int64_t)i * n + r] * y[i];
    x[r] = x[r] + (diag ? dx / diag[r] : dx);
  }
}

// ---------------------------------------------------------------------------
// The restart loop of gmres.py:214-266 on the device (esr_gmres_loop_*): the
// loop test after each true residual, the cycle's least-squares problem and
// the update, so a captured graph can run several restart cycles with no
// host round trip. Every kernel is a no-op once `done` is set.
// ---------------------------------------------------------------------------
// tol = max(rtol ||b||, atol) if atol > 0 else rtol ||b|| (gmres.py's solve)
__global__ void k_gm_loop_arm(GmresState *st, double rtol, double atol, int64_t maxiter) {
  double tol = rtol * st->bnorm;
  if (atol > 0.0 && atol > tol) tol = atol;
  st->tol = tol;
  st->maxiter = maxiter;
  st->iters = 0;
  st->done = 0;
}

// ||r|| of the residual kernel's partials, then the loop test: exit when
// ||r|| <= tol, at the cycle cap, or on a NaN residual without a cap; else
// count the cycle about to run and clear the breakdown flag. On exit the
// breakdown flag is set as well, which makes the Arnoldi kernels no-ops.
// (Every thread reads `done` before the first barrier; thread 0 writes only
// after the barriers of sum_partials.)
__global__ __launch_bounds__(kRB) void k_gm_residual_fin_loop(int64_t g,
                                                             const double *__restrict__ part,
                                                             GmresState *st) {
  if (st->done) return;
  double tot[1];
  sum_partials<1>(part, g, 0, g, tot);
  if (threadIdx.x == 0) {
    const double rn = cr_sqrt(tot[0]);
    st->rnorm = rn;
    st->jend = -1;
    const bool conv = rn == rn && rn <= st->tol;
    const bool capped = st->maxiter >= 0 && st->iters >= st->maxiter;
    const bool nan_stop = rn != rn && st->maxiter < 0;
    if (conv || capped || nan_stop) {
      st->done = 1;
      st->broken = 1;
    } else {
      st->iters += 1;
      st->broken = 0;
    }
  }
}

// min || rnorm e_0 - H y || over the cycle's jend + 1 columns (rows 0 ..
// jend + 1 of the upper Hessenberg H): Givens rotations zero the
// subdiagonal one column at a time, then back substitution. One wavefront;
// lane k holds column k. (The reference solves the same least-squares
// problem by a thin SVD, gmres.py:255-258; both give its solution up to
// rounding when H has full column rank. A zero pivot gives y_i = 0.)
__global__ __launch_bounds__(64) void k_gm_lstsq(GmresState *st) {
  if (st->done) return;
  __shared__ double R[kGmresMaxRestart + 1][kGmresMaxRestart + 1];
  __shared__ double gv[kGmresMaxRestart + 1];
  __shared__ double cs[2];
  const int k = threadIdx.x;
  const int nc = (int)st->jend + 1;
  if (nc <= 0) return;
  for (int i = 0; i <= nc; ++i)
    if (k < nc) R[i][k] = st->H[i * kGmresMaxRestart + k];
  if (k <= nc) gv[k] = k == 0 ? st->rnorm : 0.0;
  __syncthreads();
  for (int i = 0; i < nc; ++i) {
    if (k == 0) {
      const double a = R[i][i], b = R[i + 1][i];
      const double h = sqrt(a * a + b * b);
      cs[0] = h > 0.0 ? a / h : 1.0;
      cs[1] = h > 0.0 ? b / h : 0.0;
    }
    __syncthreads();
    const double c = cs[0], sn = cs[1];
    if (k >= i && k < nc) {
      const double top = R[i][k], bot = R[i + 1][k];
      R[i][k] = c * top + sn * bot;
      R[i + 1][k] = -sn * top + c * bot;
    }
    if (k == 0) {
      const double gi = gv[i];
      gv[i] = c * gi;
      gv[i + 1] = -sn * gi;
    }
    __syncthreads();
  }
  __shared__ double y[kGmresMaxRestart];
  for (int i = nc - 1; i >= 0; --i) {
    double t = (k > i && k < nc) ? R[i][k] * y[k] : 0.0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    if (k == 0) {
      const double d = R[i][i];
      y[i] = d != 0.0 ? (gv[i] - t) / d : 0.0;
    }
    __syncthreads();
  }
  if (k < nc) st->y[k] = y[k];
}

// UpdateX (gmres.py:133-136) from the device solution: x += M(sum_i V_i y_i)
// over the cycle's jend + 1 basis vectors, y rounded to the vectors' dtype
// as the host path does.
template <typename T>
__global__ __launch_bounds__(kRB) void k_gm_update_x_loop(int64_t n, const T *__restrict__ V,
                                                         const T *__restrict__ diag,
                                                         T *__restrict__ x,
                                                         const GmresState *st) {
  if (st->done) return;
  const int jc = (int)st->jend + 1;
  ESR_GRID_ROWS(r, n) {
    T dx = T(0);
    for (int i = 0; i < jc; ++i) dx = dx + V[(int64_t)i * n + r] * (T)st->y[i];
    x[r] = x[r] + (diag ? dx / diag[r] : dx);
  }
}

// ---------------------------------------------------------------------------
// Mixed-precision iterative refinement (FP64 outer loop around FP32 CG)
// ---------------------------------------------------------------------------
struct IrState {
  double rnorm;
  double bnorm;
};

// r = b - A x in float64; writes r rounded to float32 and partials of r.r
// and b.b; k_ir_fin finishes ||r|| (and, with first != 0, ||b||).
__global__ __launch_bounds__(kRB) void k_ir_residual(
    int64_t n, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const double *__restrict__ Ae, const double *__restrict__ Av,
    const double *__restrict__ bvec, const double *__restrict__ x, float *__restrict__ r32,
    double *__restrict__ part) {
  double acc[2] = {0.0, 0.0};
  ESR_GRID_ROWS(i, n) {
    const double ri =
        bvec[i] - linsys_row<double>(i, rowptr, col, Ae, Av, [&](int64_t c) { return x[c]; });
    r32[i] = (float)ri;
    acc[0] += ri * ri;
    acc[1] += bvec[i] * bvec[i];
  }
  store_partials<2>(acc, part);
}

__global__ __launch_bounds__(kRB) void k_ir_fin(int64_t g, const double *__restrict__ part,
                                               IrState *st, int first) {
  double tot[2];
  sum_partials<2>(part, g, 0, g, tot);
  if (threadIdx.x == 0) {
    st->rnorm = cr_sqrt(tot[0]);
    if (first) st->bnorm = cr_sqrt(tot[1]);
  }
}

// x += (double) d; d = 0 (the next inner solve starts from zero).
__global__ __launch_bounds__(kRB) void k_ir_update(int64_t n, double *__restrict__ x,
                                                  float *__restrict__ d) {
  ESR_GRID_ROWS(i, n) {
    x[i] = x[i] + (double)d[i];
    d[i] = 0.0f;
  }
}

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

// Workspace of the PCG kernels: partials of Init (2 G), Step kernel 1 (G)
// and kernel 2 (2 G), G = red_grid(n) workgroups -- kernel 1's and kernel
// 2's in different regions, so a kernel never overwrites what it is still
// summing -- then, 256-byte aligned, the second direction buffer (n values
// of up to 8 bytes).
// Largest system whose PCG iteration is two kernels: below it the
// iteration is launch-bound and one kernel fewer pays; above it the doubled
// neighbour gathers of the fused Update cost more than a separate pass over
// z and p (measured at 10,201 vertices: 11.8 -> 10.8 us per iteration; at
// 4.2 M: 170 -> 180 us).
constexpr int64_t kFuseUpdateMax = 1 << 17;

// Grid of the product kernel of the three-kernel iteration (its partials
// are summed by the next kernel in every workgroup, so it may differ from
// the other kernels' grid). At 4.2 M vertices caps of 1,024 / 4,096 / 8,192
// workgroups gave 158.0 / 159.4 / 160.2 us per iteration (float64) and
// 97.4 / 98.6 / 99.5 us (float32): more rows in flight do not help, the
// product is bound by the number of load instructions per row, which the
// wave-uniform slot count (linsys_row_wave) cut from 8 to 6 per row on the
// 2-triangle mesh (167.6 -> 158.0 us, 107.3 -> 97.4 us).
static unsigned spmv_grid(int64_t n) { return red_grid(n); }

static int64_t pcg_parts_bytes(int64_t n) {
  const int64_t g = red_grid(n), g1 = spmv_grid(n);
  return ((2 * g + (g1 > g ? g1 : g)) * (int64_t)sizeof(double) + 255) / 256 * 256;
}

ESR_API int esr_pcg_workspace(int64_t n, int64_t *state_bytes, int64_t *ws_bytes) {
  ESR_REQUIRE(n >= 1 && state_bytes && ws_bytes, "esr_pcg_workspace: bad args");
  *state_bytes = (int64_t)sizeof(PcgState);
  *ws_bytes = pcg_parts_bytes(n) + 8 * n;
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
  // parameters go in with a plain copy; the kernels fill the rest
  ESR_HIP(hipMemcpyAsync(state, &h, sizeof(h), hipMemcpyHostToDevice, s));
  ESR_HIP(hipStreamSynchronize(s));
  const unsigned g = red_grid(n);
  ESR_FLOAT_DISPATCH(dtype, (k_pcg_init<T><<<g, kRB, 0, s>>>(
                                n, rowptr, col, (const T *)Ae, (const T *)Av, (const T *)diag,
                                (const T *)b, (const T *)x, (T *)r, (T *)z, (T *)p,
                                (double *)ws)));
  k_pcg_init_fin<<<1, kRB, 0, s>>>(g, (const double *)ws, (PcgState *)state);
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
  const unsigned g = red_grid(n);
  const bool fuse = n <= kFuseUpdateMax;
  const unsigned g1 = fuse ? g : spmv_grid(n);
  double *part2 = (double *)ws, *part1 = part2 + 2 * (int64_t)g;
  void *p1 = (char *)ws + pcg_parts_bytes(n);
  for (int64_t it = 0; it < count; ++it) {
    ESR_FLOAT_DISPATCH(dtype, {
      if (fuse)
        k_pcg_step1<T, true><<<g1, kRB, 0, s>>>(n, rowptr, col, (const T *)Ae, (const T *)Av,
                                               (const T *)z, (T *)p, (T *)p1, (T *)Ap,
                                               (PcgState *)state, part2, part1);
      else
        k_pcg_step1<T, false><<<g1, kRB, 0, s>>>(n, rowptr, col, (const T *)Ae, (const T *)Av,
                                                (const T *)z, (T *)p, (T *)p1, (T *)Ap,
                                                (PcgState *)state, part2, part1);
      k_pcg_step2<T><<<g, kRB, 0, s>>>(n, (const T *)diag, (T *)x, (T *)r, (T *)z,
                                       (const T *)p, (const T *)p1, (const T *)Ap,
                                       (PcgState *)state, part1, (int64_t)g1, part2);
      if (!fuse)
        k_pcg_update<T><<<g, kRB, 0, s>>>(n, (const T *)z, (T *)p, (T *)p1, (PcgState *)state,
                                          part2);
    });
    ESR_LAUNCHED();
  }
  if (count > 0 && fuse) {
    ESR_FLOAT_DISPATCH(dtype, (k_pcg_settle<T><<<g, kRB, 0, s>>>(n, (T *)p, (const T *)p1,
                                                                 (PcgState *)state, part2)));
    ESR_LAUNCHED();
  }
  return ESR_OK;
  ESR_GUARD_END
}

// esr_pcg_steps with the product on the operator's SELL-64 form (built by
// esr_sell_plan / esr_sell_fill from the same CSR): the same kernels and
// results otherwise.
ESR_API int esr_pcg_steps_sell(int dtype, int64_t n, const int64_t *soff, const uint8_t *cnt,
                               const void *colv, int32_t idx16, const void *val,
                               const void *Av, const void *diag, void *x, void *r, void *z,
                               void *p, void *Ap, void *state, void *ws, int64_t count,
                               void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1 && count >= 0 && soff && cnt && colv && val && Av && x && r && z && p &&
                  Ap && state && ws,
              "esr_pcg_steps_sell: null argument");
  hipStream_t s = as_stream(stream);
  const unsigned g = red_grid(n);
  const bool fuse = n <= kFuseUpdateMax;
  const unsigned g1 = fuse ? g : spmv_grid(n);
  double *part2 = (double *)ws, *part1 = part2 + 2 * (int64_t)g;
  void *p1 = (char *)ws + pcg_parts_bytes(n);
#define ESR_STEP1_SELL(C, F)                                                           \
  k_pcg_step1_sell<T, C, F><<<g1, kRB, 0, s>>>(                                        \
      n, soff, cnt, (const C *)colv, (const T *)val, (const T *)Av, (const T *)z, (T *)p, \
      (T *)p1, (T *)Ap, (PcgState *)state, part2, part1)
  for (int64_t it = 0; it < count; ++it) {
    ESR_FLOAT_DISPATCH(dtype, {
      if (fuse) {
        if (idx16) ESR_STEP1_SELL(int16_t, true); else ESR_STEP1_SELL(int32_t, true);
      } else {
        if (idx16) ESR_STEP1_SELL(int16_t, false); else ESR_STEP1_SELL(int32_t, false);
      }
      k_pcg_step2<T><<<g, kRB, 0, s>>>(n, (const T *)diag, (T *)x, (T *)r, (T *)z,
                                       (const T *)p, (const T *)p1, (const T *)Ap,
                                       (PcgState *)state, part1, (int64_t)g1, part2);
      if (!fuse)
        k_pcg_update<T><<<g, kRB, 0, s>>>(n, (const T *)z, (T *)p, (T *)p1, (PcgState *)state,
                                          part2);
    });
    ESR_LAUNCHED();
  }
#undef ESR_STEP1_SELL
  if (count > 0 && fuse) {
    ESR_FLOAT_DISPATCH(dtype, (k_pcg_settle<T><<<g, kRB, 0, s>>>(n, (T *)p, (const T *)p1,
                                                                 (PcgState *)state, part2)));
    ESR_LAUNCHED();
  }
  return ESR_OK;
  ESR_GUARD_END
}

// Continue a stopped PCG after the caller replaced r by a freshly computed
// residual (a reliable update): z = M r, beta = r.z / rz, p = z + beta p,
// new absolute exit tolerance `tol`; the step counter continues.
ESR_API int esr_pcg_rearm(int dtype, int64_t n, const void *diag, const void *r, void *z,
                          void *p, void *state, void *ws, double tol, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1 && r && z && p && state && ws, "esr_pcg_rearm: null argument");
  hipStream_t s = as_stream(stream);
  const unsigned g = red_grid(n);
  ESR_FLOAT_DISPATCH(dtype, {
    k_pcg_rearm<T><<<g, kRB, 0, s>>>(n, (const T *)diag, (const T *)r, (T *)z, (double *)ws);
    k_pcg_rearm_fin<<<1, kRB, 0, s>>>(g, (const double *)ws, (PcgState *)state, tol);
    k_pcg_rearm_p<T><<<g, kRB, 0, s>>>(n, (const T *)z, (T *)p,
                                       (T *)((char *)ws + pcg_parts_bytes(n)),
                                       (const PcgState *)state);
  });
  ESR_LAUNCHED();
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

// Workspace of the GMRES kernels: one region of red_grid(n) partials (the
// norms of b and r) and two regions of kGmresMaxRestart rows of gs_grid(n)
// partials, used alternately by the Gram-Schmidt reductions.
static int64_t gmres_ws_doubles(int64_t n) {
  return std::max<int64_t>((int64_t)red_grid(n), 2 * (int64_t)kGmresMaxRestart * gs_grid(n));
}

ESR_API int esr_gmres_workspace(int64_t n, int64_t restart, int64_t *state_bytes,
                                int64_t *ws_bytes) {
  ESR_REQUIRE(n >= 1 && restart >= 1 && state_bytes && ws_bytes, "esr_gmres_workspace: bad");
  ESR_REQUIRE(restart <= kGmresMaxRestart, "esr_gmres_workspace: restart %lld > %d",
              (long long)restart, kGmresMaxRestart);
  *state_bytes = (int64_t)sizeof(GmresState);
  *ws_bytes = gmres_ws_doubles(n) * (int64_t)sizeof(double);
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
  const unsigned g = red_grid(n);
  ESR_FLOAT_DISPATCH(dtype, (k_gm_bnorm<T><<<g, kRB, 0, s>>>(n, (const T *)b, (double *)ws)));
  k_gm_bnorm_fin<<<1, kRB, 0, s>>>(g, (const double *)ws, (GmresState *)state);
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

// r = b - A x and ||r|| (a new cycle starts: breakdown flag cleared).
ESR_API int esr_gmres_residual(int dtype, int64_t n, const int32_t *rowptr, const int32_t *col,
                               const void *Ae, const void *Av, const void *b, const void *x,
                               void *r, void *state, void *ws, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1 && rowptr && col && Ae && Av && b && x && r && state && ws,
              "esr_gmres_residual: null argument");
  hipStream_t s = as_stream(stream);
  const unsigned g = red_grid(n);
  ESR_FLOAT_DISPATCH(dtype, (k_gm_residual<T><<<g, kRB, 0, s>>>(
                                n, rowptr, col, (const T *)Ae, (const T *)Av, (const T *)b,
                                (const T *)x, (T *)r, (double *)ws, nullptr)));
  k_gm_residual_fin<<<1, kRB, 0, s>>>(g, (const double *)ws, (GmresState *)state);
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

// One Arnoldi cycle after esr_gmres_residual: V (restart x n, one basis
// vector per row), w0/w1 two n-vectors used alternately; every kernel is a
// no-op after a breakdown. The host reads ||r||, H, jend and the breakdown
// flag afterwards (esr_gmres_read).
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
  const unsigned g = red_grid(n), gg = gs_grid(n);
  // launch-bound systems: the product and the first dots in one kernel
  const bool fuse = n <= kFuseUpdateMax;
  double *reg[2] = {(double *)ws, (double *)ws + (int64_t)kGmresMaxRestart * gg};
  ESR_FLOAT_DISPATCH(dtype, {
    T *Vt = (T *)V;
    T *w[2] = {(T *)w0, (T *)w1};
    GmresState *st = (GmresState *)state;
    for (int j = 0; j < (int)restart; ++j) {
      const T *src = j == 0 ? (const T *)r : w[(j - 1) & 1];
      T *wj = w[j & 1];
      // the norm partials of w_{j-1} are in reg[0] (written by k_gm_cgs<0>)
      if (!fuse)
        k_gm_spmv<T><<<g, kRB, 0, s>>>(n, j, rowptr, col, (const T *)Ae, (const T *)Av,
                                       (const T *)diag, src, Vt + (int64_t)j * n, wj, st,
                                       reg[0], (int64_t)gg);
      auto orth = [&](auto jb) {
        constexpr int JB = decltype(jb)::value;
        if (fuse)
          k_gm_spmv_dots<T, JB><<<gg, kRB, 0, s>>>(n, j, rowptr, col, (const T *)Ae,
                                                   (const T *)Av, (const T *)diag, src, Vt, wj,
                                                   st, reg[0], (int64_t)gg, reg[1]);
        else
          k_gm_dots<T, JB><<<gg, kRB, 0, s>>>(n, j, Vt, wj, st, reg[1]);
        k_gm_cgs<T, JB, true><<<gg, kRB, 0, s>>>(n, j, Vt, wj, st, reg[1], reg[0]);
        k_gm_cgs<T, JB, false><<<gg, kRB, 0, s>>>(n, j, Vt, wj, st, reg[0], reg[1]);
      };
      if (j < 8) orth(std::integral_constant<int, 8>());
      else if (j < 16) orth(std::integral_constant<int, 16>());
      else orth(std::integral_constant<int, 32>());
      std::swap(reg[0], reg[1]);   // the norm partials now in reg[0]
    }
    k_gm_fin<T><<<1, kRB, 0, s>>>(gg, (int)restart, st, reg[0]);
  });
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

// The restart loop on the device: esr_gmres_loop_arm after esr_gmres_begin
// (tolerance from ||b||, cycle cap, counters); each esr_gmres_loop_cycles call
// enqueues `count` restart cycles -- true residual and loop test, Arnoldi
// cycle, least squares, update of x -- every kernel a no-op once the loop
// has exited (safe to capture and replay; the host reads `done`, `iters`
// and ||r|| with esr_gmres_loop_read).
ESR_API int esr_gmres_loop_arm(void *state, double rtol, double atol, int64_t maxiter,
                               void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(state, "esr_gmres_loop_arm: null state");
  k_gm_loop_arm<<<1, 1, 0, as_stream(stream)>>>((GmresState *)state, rtol, atol, maxiter);
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

ESR_API int esr_gmres_loop_cycles(int dtype, int64_t n, int64_t restart, const int32_t *rowptr,
                                  const int32_t *col, const void *Ae, const void *Av,
                                  const void *diag, const void *b, void *x, void *r, void *V,
                                  void *w0, void *w1, void *state, void *ws, int64_t count,
                                  void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1 && restart >= 1 && restart <= kGmresMaxRestart && count >= 0,
              "esr_gmres_loop_cycles: bad sizes");
  ESR_REQUIRE(rowptr && col && Ae && Av && b && x && r && V && w0 && w1 && state && ws,
              "esr_gmres_loop_cycles: null argument");
  hipStream_t s = as_stream(stream);
  const unsigned g = red_grid(n);
  GmresState *st = (GmresState *)state;
  for (int64_t c = 0; c < count; ++c) {
    ESR_FLOAT_DISPATCH(dtype, (k_gm_residual<T><<<g, kRB, 0, s>>>(
                                  n, rowptr, col, (const T *)Ae, (const T *)Av, (const T *)b,
                                  (const T *)x, (T *)r, (double *)ws, st)));
    k_gm_residual_fin_loop<<<1, kRB, 0, s>>>(g, (const double *)ws, st);
    ESR_LAUNCHED();
    const int rc = esr_gmres_cycle(dtype, n, restart, rowptr, col, Ae, Av, diag, r, V, w0, w1,
                                   state, ws, stream);
    if (rc != ESR_OK) return rc;
    k_gm_lstsq<<<1, 64, 0, s>>>(st);
    ESR_FLOAT_DISPATCH(dtype, (k_gm_update_x_loop<T><<<g, kRB, 0, s>>>(
                                  n, (const T *)V, (const T *)diag, (T *)x, st)));
    ESR_LAUNCHED();
  }
  return ESR_OK;
  ESR_GUARD_END
}

// scalars = {rnorm, bnorm}, ints = {iters, done}.
ESR_API int esr_gmres_loop_read(const void *state, double *scalars, int64_t *ints,
                                void *stream) {
  ESR_REQUIRE(state && scalars && ints, "esr_gmres_loop_read: bad argument");
  GmresState h;
  hipStream_t s = as_stream(stream);
  ESR_HIP(hipMemcpyAsync(&h, state, sizeof(h), hipMemcpyDeviceToHost, s));
  ESR_HIP(hipStreamSynchronize(s));
  scalars[0] = h.rnorm;
  scalars[1] = h.bnorm;
  ints[0] = h.iters;
  ints[1] = h.done;
  return ESR_OK;
}

ESR_API int esr_gmres_update_x(int dtype, int64_t n, int64_t jcount, const void *V,
                               const void *y, const void *diag, void *x, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1 && jcount >= 0 && jcount <= kGmresMaxRestart && V && y && x,
              "esr_gmres_update_x: bad argument");
  hipStream_t s = as_stream(stream);
  ESR_FLOAT_DISPATCH(dtype, (k_gm_update_x<T><<<red_grid(n), kRB, 0, s>>>(
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
  *ws_bytes = 2 * (int64_t)red_grid(n) * (int64_t)sizeof(double);
  return ESR_OK;
}

ESR_API int esr_ir_residual(int64_t n, const int32_t *rowptr, const int32_t *col, const double *Ae,
                            const double *Av, const double *b, const double *x, float *r32,
                            void *state, void *ws, int first, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1 && rowptr && col && Ae && Av && b && x && r32 && state && ws,
              "esr_ir_residual: null argument");
  hipStream_t s = as_stream(stream);
  const unsigned g = red_grid(n);
  k_ir_residual<<<g, kRB, 0, s>>>(n, rowptr, col, Ae, Av, b, x, r32, (double *)ws);
  k_ir_fin<<<1, kRB, 0, s>>>(g, (const double *)ws, (IrState *)state, first);
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

ESR_API int esr_ir_update(int64_t n, double *x, float *d, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1 && x && d, "esr_ir_update: null argument");
  k_ir_update<<<red_grid(n), kRB, 0, as_stream(stream)>>>(n, x, d);
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
