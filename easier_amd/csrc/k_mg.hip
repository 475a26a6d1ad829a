// Geometric multigrid on the 2-triangle mesh: the hierarchy and the V-cycle
// kernels of the multilevel preconditioner used inside flexible CG.
// No reference counterpart: the reference's preconditioner slot
// (easier/numeric/preconditioner/__init__.py:4) is empty and its solvers take
// any callable M (numeric/solver/cg.py:86).
//
// Grids. Level l is the (n+1) x (n+1) vertex grid of the 2-triangle mesh
// (squares split along the SW-NE diagonal); vertex (I, J) sits at
// lexicographic position J (n+1) + I. Coarse levels are stored in
// lexicographic order; the finest level may be in any order, given by
// fpos (fine index -> position) and fidx (position -> fine index); the cycle
// kernels read a parent map built from fpos instead (esr_mg_parents).
//
// Transfer. P is P1 interpolation between the nested meshes: a fine vertex
// at an even-even position copies its coarse parent, one on a horizontal,
// vertical or SW-NE coarse edge takes the mean of the edge's two ends, and
// coarse boundary vertices (zero Dirichlet values) contribute nothing.
// Restriction is P^T, gathered per coarse vertex over its 7 children in a
// fixed order (no atomics, deterministic).
//
// Coarse operators are Galerkin products P^T A P on the coarse mesh's vertex
// graph (7-point stencil: SW, S, W, E, N, NE + the diagonal), one thread per
// coarse vertex summing its children's rows in a fixed order; boundary rows
// are identity rows, as in the P1 assembly.
#include "esr_common.h"
#include "esr_krylov.h"

namespace esr {
namespace {

constexpr int kMB = 256;

__device__ __forceinline__ int64_t mg_tid() { return (int64_t)blockIdx.x * kMB + threadIdx.x; }

// Parent map of a fine level: par[2i], par[2i+1] = the coarse indices of
// vertex i's parents, -1 for a coarse boundary vertex (zero value), and
// par[2i+1] = -2 when i sits on a coarse vertex (one parent, weight 1; else
// two parents, weight 1/2 each). Built once per hierarchy, so the cycle
// kernels do no index division.
__device__ __forceinline__ void parents_of(int64_t p, int64_t nf, int32_t &p0, int32_t &p1) {
  const uint32_t Wf = (uint32_t)(nf + 1), nc = (uint32_t)(nf >> 1), Wc = nc + 1;
  const uint32_t I = (uint32_t)p % Wf, J = (uint32_t)p / Wf;
  const uint32_t a = I >> 1, b = J >> 1;
  auto C = [&](uint32_t x, uint32_t y) -> int32_t {
    return (x > 0 && y > 0 && x < nc && y < nc) ? (int32_t)(y * Wc + x) : -1;
  };
  p0 = C(a, b);
  if (!(I & 1) && !(J & 1)) p1 = -2;
  else if (!(J & 1)) p1 = C(a + 1, b);
  else if (!(I & 1)) p1 = C(a, b + 1);
  else p1 = C(a + 1, b + 1);
}

// (P zc) at fine vertex i: the parent's value, or 0.5 (first + second).
template <typename T>
__device__ __forceinline__ T prolong_value(int64_t i, const int32_t *__restrict__ par,
                                           const T *__restrict__ zc) {
  const int2 pp = reinterpret_cast<const int2 *>(par)[i];
  const T v0 = pp.x >= 0 ? zc[pp.x] : T(0);
  if (pp.y == -2) return v0;
  const T v1 = pp.y >= 0 ? zc[pp.y] : T(0);
  return T(0.5) * (v0 + v1);
}

// ---------------------------------------------------------------------------
// set-up
// ---------------------------------------------------------------------------
// Vertex graph of the n-grid in lexicographic order, neighbours ascending
// (SW, S, W, E, N, NE where they exist). rowptr in closed form: a row of the
// grid holds (2W-1)([y>0] + [y<W-1]) + 2(W-1) entries.
__global__ __launch_bounds__(kMB) void k_mg_pattern(int64_t n, int32_t *__restrict__ rowptr,
                                                   int32_t *__restrict__ col) {
  const int64_t W = n + 1, v = mg_tid();
  if (v >= W * W) return;
  const int64_t x = v % W, y = v / W;
  const int64_t ym = y > 0 ? y - 1 : 0, xm = x > 0 ? x - 1 : 0;
  const int64_t full = (2 * W - 1) * (ym + y) + 2 * (W - 1) * y;
  const int64_t part = (y > 0 ? xm + x : 0) + xm + x + (y < W - 1 ? 2 * x : 0);
  int64_t k = full + part;
  rowptr[v] = (int32_t)k;
  if (v == W * W - 1) rowptr[W * W] = (int32_t)((W - 1) * (6 * W - 2));
  const int dx[6] = {-1, 0, -1, 1, 0, 1}, dy[6] = {-1, -1, 0, 0, 1, 1};
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const int64_t xx = x + dx[s], yy = y + dy[s];
    if (xx >= 0 && xx < W && yy >= 0 && yy < W) col[k++] = (int32_t)(yy * W + xx);
  }
}

__global__ __launch_bounds__(kMB) void k_mg_parents(int64_t nf, const int32_t *__restrict__ fpos,
                                                   int32_t *__restrict__ par) {
  const int64_t nv = (nf + 1) * (nf + 1), i = mg_tid();
  if (i >= nv) return;
  int32_t p0, p1;
  parents_of(fpos ? fpos[i] : i, nf, p0, p1);
  reinterpret_cast<int2 *>(par)[i] = make_int2(p0, p1);
}

// Slot of coarse offset (dx, dy): 0 = diagonal, 1..6 = SW, S, W, E, N, NE;
// -1 outside the stencil.
__device__ __forceinline__ int stencil_slot(int64_t dx, int64_t dy) {
  if (dx == 0 && dy == 0) return 0;
  if (dx == -1 && dy == -1) return 1;
  if (dx == 0 && dy == -1) return 2;
  if (dx == -1 && dy == 0) return 3;
  if (dx == 1 && dy == 0) return 4;
  if (dx == 0 && dy == 1) return 5;
  if (dx == 1 && dy == 1) return 6;
  return -1;
}

__global__ __launch_bounds__(kMB) void k_mg_galerkin(
    int64_t nf, const int32_t *__restrict__ fidx, const int32_t *__restrict__ fpos,
    const int32_t *__restrict__ f_rowptr, const int32_t *__restrict__ f_col,
    const double *__restrict__ f_Ae, const double *__restrict__ f_Av,
    const int32_t *__restrict__ c_rowptr, double *__restrict__ c_Ae, double *__restrict__ c_Av,
    int32_t *__restrict__ bad) {
  const int64_t nc = nf >> 1, Wc = nc + 1, Wf = nf + 1, c = mg_tid();
  if (c >= Wc * Wc) return;
  const int64_t a = c % Wc, b = c / Wc;
  const int32_t cb = c_rowptr[c], ce = c_rowptr[c + 1];
  if (a == 0 || b == 0 || a == nc || b == nc) {
    c_Av[c] = 1.0;
    for (int32_t e = cb; e < ce; ++e) c_Ae[e] = 0.0;
    return;
  }
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  // children in the restriction's order: self, W, E, S, N, SW, NE
  const int cdx[7] = {0, -1, 1, 0, 0, -1, 1}, cdy[7] = {0, 0, 0, -1, 1, -1, 1};
  int nbad = 0;
  for (int k = 0; k < 7; ++k) {
    const double wk = k == 0 ? 1.0 : 0.5;
    const int64_t p = (2 * b + cdy[k]) * Wf + 2 * a + cdx[k];
    const int64_t i = fidx ? fidx[p] : p;
    const int32_t rb = f_rowptr[i], re = f_rowptr[i + 1];
    for (int32_t e = rb - 1; e < re; ++e) {   // e = rb - 1: the diagonal entry
      const int64_t j = e < rb ? i : f_col[e];
      const double v = wk * (e < rb ? f_Av[i] : f_Ae[e]);
      if (v == 0.0) continue;
      const int64_t q = fpos ? fpos[j] : j;
      const int64_t I = q % Wf, J = q / Wf, pa = I >> 1, pb = J >> 1;
      // parents of j (weights 1 or 1/2), interior ones only
      int64_t ax[2], by[2];
      double w;
      int np = 2;
      ax[0] = pa; by[0] = pb;
      if (!(I & 1) && !(J & 1)) { np = 1; w = 1.0; }
      else if (!(J & 1)) { ax[1] = pa + 1; by[1] = pb; w = 0.5; }
      else if (!(I & 1)) { ax[1] = pa; by[1] = pb + 1; w = 0.5; }
      else { ax[1] = pa + 1; by[1] = pb + 1; w = 0.5; }
      for (int s = 0; s < np; ++s) {
        if (ax[s] <= 0 || by[s] <= 0 || ax[s] >= nc || by[s] >= nc) continue;
        const int slot = stencil_slot(ax[s] - a, by[s] - b);
        if (slot < 0) { ++nbad; continue; }
        acc[slot] += v * w;
      }
    }
  }
  if (nbad) atomicAdd(bad, nbad);
  c_Av[c] = acc[0];
  // interior coarse vertices hold all six neighbours, ascending
  for (int s = 0; s < 6 && cb + s < ce; ++s) c_Ae[cb + s] = acc[s + 1];
}

// Network inputs of a lexicographic level (the learned smoother):
// vfeat = (log(d / mean d), log kappa), efeat = (a_ij / sqrt(d_i d_j), dx, dy)
// with (dx, dy) the neighbour's offset in grid units.
__global__ __launch_bounds__(kMB) void k_mg_features(int64_t n, const int32_t *__restrict__ rowptr,
                                                    const int32_t *__restrict__ col,
                                                    const double *__restrict__ Ae,
                                                    const double *__restrict__ Av,
                                                    const double *__restrict__ dsum,
                                                    const float *__restrict__ logk,
                                                    float *__restrict__ vfeat,
                                                    float *__restrict__ efeat) {
  const int64_t W = n + 1, nv = W * W, i = mg_tid();
  if (i >= nv) return;
  const double dmean = dsum[0] / (double)nv;
  const double di = Av[i];
  vfeat[2 * i] = (float)log(di / dmean);
  vfeat[2 * i + 1] = logk[i];
  const int64_t x = i % W, y = i / W;
  for (int32_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
    const int64_t j = col[e];
    efeat[3 * (int64_t)e] = (float)(Ae[e] / sqrt(di * Av[j]));
    efeat[3 * (int64_t)e + 1] = (float)(j % W - x);
    efeat[3 * (int64_t)e + 2] = (float)(j / W - y);
  }
}

// ---------------------------------------------------------------------------
// V-cycle
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kMB) void k_mg_restrict(int64_t nf, const int32_t *__restrict__ fidx,
                                                    const T *__restrict__ t, T *__restrict__ rc) {
  const int64_t nc = nf >> 1, Wc = nc + 1, Wf = nf + 1, c = mg_tid();
  if (c >= Wc * Wc) return;
  const int64_t a = (uint32_t)c % (uint32_t)Wc, b = (uint32_t)c / (uint32_t)Wc;
  if (a == 0 || b == 0 || a == nc || b == nc) {
    rc[c] = T(0);
    return;
  }
  const int64_t p0 = 2 * b * Wf + 2 * a;
  auto F = [&](int64_t p) -> T { return t[fidx ? fidx[p] : p]; };
  T acc = F(p0 - 1);
  acc = acc + F(p0 + 1);
  acc = acc + F(p0 - Wf);
  acc = acc + F(p0 + Wf);
  acc = acc + F(p0 - Wf - 1);
  acc = acc + F(p0 + Wf + 1);
  rc[c] = F(p0) + T(0.5) * acc;
}

// z += P zc
template <typename T>
__global__ __launch_bounds__(kMB) void k_mg_prolong(int64_t nf, const int32_t *__restrict__ par,
                                                   const T *__restrict__ zc, T *__restrict__ z) {
  const int64_t nv = (nf + 1) * (nf + 1), i = mg_tid();
  if (i >= nv) return;
  z[i] = z[i] + prolong_value<T>(i, par, zc);
}

// Damped-Jacobi pre-smoothing from z = 0 and the residual it leaves:
// z = wd r; t = r - A z, with wd = w / d per row (precomputed, so the sweep
// multiplies instead of dividing at every neighbour).
template <typename T, typename Op>
__global__ __launch_bounds__(kMB) void k_mg_jacobi_pre(int64_t n, const Op A,
                                                      const T *__restrict__ wd,
                                                      const T *__restrict__ r, T *__restrict__ z,
                                                      T *__restrict__ t) {
  const int64_t i = mg_tid();
  if (i >= n) return;
  const T az = A.row(i, [&](int64_t j) { return wd[j] * r[j]; });
  z[i] = wd[i] * r[i];
  t[i] = r[i] - az;
}

// Coarse-grid correction and damped-Jacobi post-smoothing:
// y = z + P zc; out = y + wd (r - A y), wd = w / d. With zc null, y = z (the
// correction already added by k_mg_prolong): the fused form evaluates P at
// every neighbour (a rowptr -> col -> parent -> zc load chain), which runs
// about 3x slower on the finest level than the two-pass form.
template <typename T, typename Op>
__global__ __launch_bounds__(kMB) void k_mg_jacobi_post(int64_t nv, const int32_t *__restrict__ par,
                                                       const Op A, const T *__restrict__ wd,
                                                       const T *__restrict__ r,
                                                       const T *__restrict__ z,
                                                       const T *__restrict__ zc,
                                                       T *__restrict__ out) {
  const int64_t i = mg_tid();
  if (i >= nv) return;
  if (zc) {
    auto y = [&](int64_t j) { return z[j] + prolong_value<T>(j, par, zc); };
    const T ay = A.row(i, y);
    const T yi = y(i);
    out[i] = yi + wd[i] * (r[i] - ay);
  } else {
    const T ay = A.row(i, [&](int64_t j) { return z[j]; });
    out[i] = z[i] + wd[i] * (r[i] - ay);
  }
}

// t = r - A y with y = z (zc null) or y = z + P zc (then also stored).
template <typename T, typename Op>
__global__ __launch_bounds__(kMB) void k_mg_residual(int64_t n, int64_t nf,
                                                    const int32_t *__restrict__ par, const Op A,
                                                    const T *__restrict__ r,
                                                    const T *__restrict__ z,
                                                    const T *__restrict__ zc, T *__restrict__ y,
                                                    T *__restrict__ t) {
  const int64_t i = mg_tid();
  if (i >= n) return;
  if (zc) {
    auto yv = [&](int64_t j) { return z[j] + prolong_value<T>(j, par, zc); };
    const T ay = A.row(i, yv);
    y[i] = yv(i);
    t[i] = r[i] - ay;
  } else {
    const T az = A.row(i, [&](int64_t j) { return z[j]; });
    t[i] = r[i] - az;
  }
}

// Coarsest grid: z = Ainv r (dense, row-major), one wavefront per row.
template <typename T>
__global__ __launch_bounds__(kMB) void k_mg_dense(int64_t n, const double *__restrict__ Ainv,
                                                 const T *__restrict__ r, T *__restrict__ z) {
  const int64_t row = mg_tid() >> 6;
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  double s = 0.0;
  for (int64_t j = lane; j < n; j += 64) s += Ainv[row * n + j] * (double)r[j];
  s = kr_wave_sum(s);
  if (lane == 0) z[row] = (T)s;
}

// ---------------------------------------------------------------------------
// Algebraic levels (smoothed aggregation, set up on the host): the transfers
// are general CSR operators, P by fine row and R = P^T by coarse row, and
// the coarse operators general CSR matrices with rows of 10-60 entries.
// A thread per row would walk such a row as one chain of dependent loads
// (about 250 us for the 700 k rows of ~20 entries that restrict a 4.2 M-row
// level), so rows are spread over L lanes (L a power of two <= 64, the L
// lanes of a row inside one wavefront): lane l folds entries l, l + L, ...
// in order from 0, and the L partial sums are combined by an xor butterfly,
// whose sums are bitwise the same on every lane (IEEE addition commutes).
// ---------------------------------------------------------------------------
template <int L, typename T, typename Src>
__device__ __forceinline__ T csr_row_lanes(int32_t b, int32_t e, int lane,
                                           const int32_t *__restrict__ col,
                                           const T *__restrict__ val, Src src) {
  T s = T(0);
  for (int32_t k = b + lane; k < e; k += L) s = s + val[k] * src(col[k]);
#pragma unroll
  for (int off = L / 2; off > 0; off >>= 1) s = s + __shfl_xor(s, off, L);
  return s;
}

// y = R t (add 0) or z += P zc (add 1), L lanes per row.
template <typename T, int L>
__global__ __launch_bounds__(kMB) void k_amg_transfer(int64_t m, const int32_t *__restrict__ rowptr,
                                                     const int32_t *__restrict__ col,
                                                     const T *__restrict__ val,
                                                     const T *__restrict__ x, T *__restrict__ y,
                                                     int add) {
  const int64_t t = mg_tid(), i = t / L;
  const int lane = (int)(t & (L - 1));
  if (i >= m) return;
  const T s = csr_row_lanes<L, T>(rowptr[i], rowptr[i + 1], lane, col, val,
                                  [&](int64_t j) { return x[j]; });
  if (lane == 0) y[i] = add ? y[i] + s : s;
}

// Damped-Jacobi sweeps of a CSR level, L lanes per row; the row product is
// the Linsys fold (off-diagonal entries, then the diagonal term):
//   mode 0 (pre):      z = wd r; t = r - A z
//   mode 1 (post):     out = z + wd (r - A z)
//   mode 2 (residual): t = r - A z
template <typename T, int L, int MODE>
__global__ __launch_bounds__(kMB) void k_amg_sweep(int64_t n, const int32_t *__restrict__ rowptr,
                                                  const int32_t *__restrict__ col,
                                                  const T *__restrict__ Ae,
                                                  const T *__restrict__ Av,
                                                  const T *__restrict__ wd,
                                                  const T *__restrict__ r, T *__restrict__ z,
                                                  T *__restrict__ out) {
  const int64_t t = mg_tid(), i = t / L;
  const int lane = (int)(t & (L - 1));
  if (i >= n) return;
  auto src = [&](int64_t j) -> T { return MODE == 0 ? wd[j] * r[j] : z[j]; };
  const T s = csr_row_lanes<L, T>(rowptr[i], rowptr[i + 1], lane, col, Ae, src);
  if (lane != 0) return;
  const T az = s + src(i) * Av[i];
  if (MODE == 0) {
    z[i] = wd[i] * r[i];
    out[i] = r[i] - az;
  } else if (MODE == 1) {
    out[i] = z[i] + wd[i] * (r[i] - az);
  } else {
    out[i] = r[i] - az;
  }
}

inline unsigned mg_grid(int64_t n) { return (unsigned)ceil_div(n, kMB); }

} // namespace
} // namespace esr

using namespace esr;

#define MG_DTYPE_DISPATCH(dtype, CALL)                                         \
  do {                                                                         \
    if ((dtype) == ESR_F64) { using T = double; CALL; }                        \
    else if ((dtype) == ESR_F32) { using T = float; CALL; }                    \
    else { set_error("multigrid: dtype %d is not float32/float64", (int)(dtype)); return ESR_ERR_UNSUPPORTED; } \
  } while (0)

ESR_API int esr_mg_pattern(int64_t n, int32_t *rowptr, int32_t *col, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1 && (n + 1) * (n + 1) * 6 < (int64_t)INT32_MAX, "esr_mg_pattern: n %lld out of range", (long long)n);
  ESR_REQUIRE(rowptr && col, "esr_mg_pattern: null pointer");
  const int64_t nv = (n + 1) * (n + 1);
  k_mg_pattern<<<mg_grid(nv), kMB, 0, as_stream(stream)>>>(n, rowptr, col);
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

ESR_API int esr_mg_galerkin(int64_t nf, const int32_t *fidx, const int32_t *fpos,
                            const int32_t *f_rowptr, const int32_t *f_col, const double *f_Ae,
                            const double *f_Av, const int32_t *c_rowptr, double *c_Ae,
                            double *c_Av, int32_t *bad, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(nf >= 2 && nf % 2 == 0, "esr_mg_galerkin: fine grid n=%lld is not even", (long long)nf);
  ESR_REQUIRE((fidx == nullptr) == (fpos == nullptr), "esr_mg_galerkin: give both fidx and fpos or neither");
  ESR_REQUIRE(f_rowptr && f_col && f_Ae && f_Av && c_rowptr && c_Ae && c_Av && bad,
              "esr_mg_galerkin: null pointer");
  const int64_t Wc = nf / 2 + 1;
  k_mg_galerkin<<<mg_grid(Wc * Wc), kMB, 0, as_stream(stream)>>>(
      nf, fidx, fpos, f_rowptr, f_col, f_Ae, f_Av, c_rowptr, c_Ae, c_Av, bad);
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

ESR_API int esr_mg_features(int64_t n, const int32_t *rowptr, const int32_t *col,
                            const double *Ae, const double *Av, const double *dsum,
                            const float *logk, float *vfeat, float *efeat, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1, "esr_mg_features: n must be >= 1");
  ESR_REQUIRE(rowptr && col && Ae && Av && dsum && logk && vfeat && efeat,
              "esr_mg_features: null pointer");
  const int64_t nv = (n + 1) * (n + 1);
  k_mg_features<<<mg_grid(nv), kMB, 0, as_stream(stream)>>>(n, rowptr, col, Ae, Av, dsum, logk,
                                                            vfeat, efeat);
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

ESR_API int esr_mg_restrict(int dtype, int64_t nf, const int32_t *fidx, const void *t, void *rc,
                            void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(nf >= 2 && nf % 2 == 0, "esr_mg_restrict: fine grid n=%lld is not even", (long long)nf);
  ESR_REQUIRE(t && rc, "esr_mg_restrict: null pointer");
  const int64_t Wc = nf / 2 + 1;
  MG_DTYPE_DISPATCH(dtype, (k_mg_restrict<T><<<mg_grid(Wc * Wc), kMB, 0, as_stream(stream)>>>(
                               nf, fidx, (const T *)t, (T *)rc)));
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

ESR_API int esr_mg_parents(int64_t nf, const int32_t *fpos, int32_t *par, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(nf >= 2 && nf % 2 == 0, "esr_mg_parents: fine grid n=%lld is not even", (long long)nf);
  ESR_REQUIRE(par, "esr_mg_parents: null pointer");
  const int64_t nv = (nf + 1) * (nf + 1);
  k_mg_parents<<<mg_grid(nv), kMB, 0, as_stream(stream)>>>(nf, fpos, par);
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

ESR_API int esr_mg_prolong(int dtype, int64_t nf, const int32_t *par, const void *zc, void *z,
                           void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(nf >= 2 && nf % 2 == 0, "esr_mg_prolong: fine grid n=%lld is not even", (long long)nf);
  ESR_REQUIRE(par && zc && z, "esr_mg_prolong: null pointer");
  const int64_t nv = (nf + 1) * (nf + 1);
  MG_DTYPE_DISPATCH(dtype, (k_mg_prolong<T><<<mg_grid(nv), kMB, 0, as_stream(stream)>>>(
                               nf, par, (const T *)zc, (T *)z)));
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

namespace {
template <typename T, typename Op>
int mg_pre(int64_t n, const Op &A, const void *wd, const void *r, void *z, void *t,
           hipStream_t s) {
  k_mg_jacobi_pre<T><<<mg_grid(n), kMB, 0, s>>>(n, A, (const T *)wd, (const T *)r, (T *)z,
                                                (T *)t);
  ESR_LAUNCHED();
  return ESR_OK;
}
template <typename T, typename Op>
int mg_post(int64_t nv, const int32_t *par, const Op &A, const void *wd, const void *r,
            const void *z, const void *zc, void *out, hipStream_t s) {
  k_mg_jacobi_post<T><<<mg_grid(nv), kMB, 0, s>>>(nv, par, A, (const T *)wd, (const T *)r,
                                                  (const T *)z, (const T *)zc, (T *)out);
  ESR_LAUNCHED();
  return ESR_OK;
}
template <typename T, typename Op>
int mg_res(int64_t n, int64_t nf, const int32_t *par, const Op &A, const void *r, const void *z,
           const void *zc, void *y, void *t, hipStream_t s) {
  k_mg_residual<T><<<mg_grid(n), kMB, 0, s>>>(n, nf, par, A, (const T *)r, (const T *)z,
                                              (const T *)zc, (T *)y, (T *)t);
  ESR_LAUNCHED();
  return ESR_OK;
}
// the operator of a sweep: CSR, or SELL-64 with int16 / int32 columns
#define MG_OP_DISPATCH(T, CALL_WITH_A)                                                     \
  do {                                                                                     \
    if (soff == nullptr) {                                                                 \
      const CsrRows<T> A{rowptr, col, (const T *)Ae, (const T *)Av};                       \
      return CALL_WITH_A;                                                                  \
    } else if (idx16) {                                                                    \
      const SellRows<T, int16_t> A{soff, cnt, (const int16_t *)colv, (const T *)val,       \
                                   (const T *)Av};                                         \
      return CALL_WITH_A;                                                                  \
    } else {                                                                               \
      const SellRows<T, int32_t> A{soff, cnt, (const int32_t *)colv, (const T *)val,       \
                                   (const T *)Av};                                         \
      return CALL_WITH_A;                                                                  \
    }                                                                                      \
  } while (0)
} // namespace

ESR_API int esr_mg_jacobi_pre(int dtype, int64_t n, const int32_t *rowptr, const int32_t *col,
                              const void *Ae, const void *Av, const void *wd, const void *r,
                              void *z, void *t, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1, "esr_mg_jacobi_pre: n must be >= 1");
  ESR_REQUIRE(rowptr && col && Ae && Av && wd && r && z && t, "esr_mg_jacobi_pre: null pointer");
  const int64_t *soff = nullptr;
  const uint8_t *cnt = nullptr;
  const void *colv = nullptr, *val = nullptr;
  const int idx16 = 0;
  MG_DTYPE_DISPATCH(dtype, MG_OP_DISPATCH(T, (mg_pre<T>(n, A, wd, r, z, t, as_stream(stream)))));
  ESR_GUARD_END
}

ESR_API int esr_mg_jacobi_post(int dtype, int64_t nf, const int32_t *par, const int32_t *rowptr,
                               const int32_t *col, const void *Ae, const void *Av, const void *wd,
                               const void *r, const void *z, const void *zc, void *out,
                               void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(nf >= 2 && nf % 2 == 0, "esr_mg_jacobi_post: fine grid n=%lld is not even", (long long)nf);
  ESR_REQUIRE(rowptr && col && Ae && Av && wd && r && z && out && (par || !zc),
              "esr_mg_jacobi_post: null pointer");
  ESR_REQUIRE(out != z, "esr_mg_jacobi_post: out must not alias z");
  const int64_t *soff = nullptr;
  const uint8_t *cnt = nullptr;
  const void *colv = nullptr, *val = nullptr;
  const int idx16 = 0;
  const int64_t nv = (nf + 1) * (nf + 1);
  MG_DTYPE_DISPATCH(dtype, MG_OP_DISPATCH(T, (mg_post<T>(nv, par, A, wd, r, z, zc, out,
                                                         as_stream(stream)))));
  ESR_GUARD_END
}

ESR_API int esr_mg_residual(int dtype, int64_t n, int64_t nf, const int32_t *par,
                            const int32_t *rowptr, const int32_t *col, const void *Ae,
                            const void *Av, const void *r, const void *z, const void *zc,
                            void *y, void *t, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1, "esr_mg_residual: n must be >= 1");
  ESR_REQUIRE(rowptr && col && Ae && Av && r && z && t, "esr_mg_residual: null pointer");
  if (zc) {
    ESR_REQUIRE(par && y && y != z,
                "esr_mg_residual: with zc, par and y must be given and y must not alias z");
    ESR_REQUIRE(nf >= 2 && nf % 2 == 0 && (nf + 1) * (nf + 1) == n,
                "esr_mg_residual: n=%lld is not the vertex count of an even grid nf=%lld",
                (long long)n, (long long)nf);
  }
  const int64_t *soff = nullptr;
  const uint8_t *cnt = nullptr;
  const void *colv = nullptr, *val = nullptr;
  const int idx16 = 0;
  MG_DTYPE_DISPATCH(dtype, MG_OP_DISPATCH(T, (mg_res<T>(n, nf, par, A, r, z, zc, y, t,
                                                        as_stream(stream)))));
  ESR_GUARD_END
}

// The three sweeps on the level operator's SELL-64 form (esr_sell_plan /
// esr_sell_fill, values in the sweep's dtype): the same results.
ESR_API int esr_mg_jacobi_pre_sell(int dtype, int64_t n, const int64_t *soff, const uint8_t *cnt,
                                   const void *colv, int32_t idx16, const void *val,
                                   const void *Av, const void *wd, const void *r, void *z,
                                   void *t, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1, "esr_mg_jacobi_pre_sell: n must be >= 1");
  ESR_REQUIRE(soff && cnt && colv && val && Av && wd && r && z && t,
              "esr_mg_jacobi_pre_sell: null pointer");
  const int32_t *rowptr = nullptr, *col = nullptr;
  const void *Ae = nullptr;
  MG_DTYPE_DISPATCH(dtype, MG_OP_DISPATCH(T, (mg_pre<T>(n, A, wd, r, z, t, as_stream(stream)))));
  ESR_GUARD_END
}

ESR_API int esr_mg_jacobi_post_sell(int dtype, int64_t nf, const int32_t *par,
                                    const int64_t *soff, const uint8_t *cnt, const void *colv,
                                    int32_t idx16, const void *val, const void *Av,
                                    const void *wd, const void *r, const void *z,
                                    const void *zc, void *out, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(nf >= 2 && nf % 2 == 0, "esr_mg_jacobi_post_sell: fine grid n=%lld is not even",
              (long long)nf);
  ESR_REQUIRE(soff && cnt && colv && val && Av && wd && r && z && out && (par || !zc),
              "esr_mg_jacobi_post_sell: null pointer");
  ESR_REQUIRE(out != z, "esr_mg_jacobi_post_sell: out must not alias z");
  const int32_t *rowptr = nullptr, *col = nullptr;
  const void *Ae = nullptr;
  const int64_t nv = (nf + 1) * (nf + 1);
  MG_DTYPE_DISPATCH(dtype, MG_OP_DISPATCH(T, (mg_post<T>(nv, par, A, wd, r, z, zc, out,
                                                         as_stream(stream)))));
  ESR_GUARD_END
}

ESR_API int esr_mg_residual_sell(int dtype, int64_t n, int64_t nf, const int32_t *par,
                                 const int64_t *soff, const uint8_t *cnt, const void *colv,
                                 int32_t idx16, const void *val, const void *Av, const void *r,
                                 const void *z, const void *zc, void *y, void *t, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1, "esr_mg_residual_sell: n must be >= 1");
  ESR_REQUIRE(soff && cnt && colv && val && Av && r && z && t,
              "esr_mg_residual_sell: null pointer");
  if (zc) {
    ESR_REQUIRE(par && y && y != z,
                "esr_mg_residual_sell: with zc, par and y must be given and y must not alias z");
    ESR_REQUIRE(nf >= 2 && nf % 2 == 0 && (nf + 1) * (nf + 1) == n,
                "esr_mg_residual_sell: n=%lld is not the vertex count of an even grid nf=%lld",
                (long long)n, (long long)nf);
  }
  const int32_t *rowptr = nullptr, *col = nullptr;
  const void *Ae = nullptr;
  MG_DTYPE_DISPATCH(dtype, MG_OP_DISPATCH(T, (mg_res<T>(n, nf, par, A, r, z, zc, y, t,
                                                        as_stream(stream)))));
  ESR_GUARD_END
}

// Algebraic levels: transfers and sweeps with `lanes` lanes per row
// (1, 2, 4, ..., 64).
#define AMG_LANES_DISPATCH(lanes, CALL)                                        \
  do {                                                                         \
    switch (lanes) {                                                           \
    case 1: { constexpr int L = 1; CALL; } break;                              \
    case 2: { constexpr int L = 2; CALL; } break;                              \
    case 4: { constexpr int L = 4; CALL; } break;                              \
    case 8: { constexpr int L = 8; CALL; } break;                              \
    case 16: { constexpr int L = 16; CALL; } break;                            \
    case 32: { constexpr int L = 32; CALL; } break;                            \
    case 64: { constexpr int L = 64; CALL; } break;                            \
    default: set_error("lanes %d is not a power of two <= 64", (int)(lanes));   \
      return ESR_ERR_INVALID;                                                  \
    }                                                                          \
  } while (0)

namespace {
inline unsigned amg_grid(int64_t rows, int lanes) { return mg_grid(rows * lanes); }
}

ESR_API int esr_amg_restrict(int dtype, int64_t nc, const int32_t *r_rowptr, const int32_t *r_col,
                             const void *r_val, const void *t, void *rc, int32_t lanes,
                             void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(nc >= 1, "esr_amg_restrict: nc must be >= 1");
  ESR_REQUIRE(r_rowptr && r_col && r_val && t && rc && t != rc,
              "esr_amg_restrict: null or aliased pointer");
  MG_DTYPE_DISPATCH(dtype, AMG_LANES_DISPATCH(lanes, (k_amg_transfer<T, L><<<
      amg_grid(nc, L), kMB, 0, as_stream(stream)>>>(nc, r_rowptr, r_col, (const T *)r_val,
                                                      (const T *)t, (T *)rc, 0))));
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

ESR_API int esr_amg_prolong(int dtype, int64_t nf, const int32_t *p_rowptr, const int32_t *p_col,
                            const void *p_val, const void *zc, void *z, int32_t lanes,
                            void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(nf >= 1, "esr_amg_prolong: nf must be >= 1");
  ESR_REQUIRE(p_rowptr && p_col && p_val && zc && z && zc != z,
              "esr_amg_prolong: null or aliased pointer");
  MG_DTYPE_DISPATCH(dtype, AMG_LANES_DISPATCH(lanes, (k_amg_transfer<T, L><<<
      amg_grid(nf, L), kMB, 0, as_stream(stream)>>>(nf, p_rowptr, p_col, (const T *)p_val,
                                                      (const T *)zc, (T *)z, 1))));
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

ESR_API int esr_amg_sweep(int dtype, int32_t mode, int64_t n, const int32_t *rowptr,
                          const int32_t *col, const void *Ae, const void *Av, const void *wd,
                          const void *r, void *z, void *out, int32_t lanes, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1, "esr_amg_sweep: n must be >= 1");
  ESR_REQUIRE(mode >= 0 && mode <= 2, "esr_amg_sweep: mode %d is not 0, 1 or 2", (int)mode);
  ESR_REQUIRE(rowptr && col && Ae && Av && r && z && out && (wd || mode == 2),
              "esr_amg_sweep: null pointer");
  ESR_REQUIRE(out != z && out != r, "esr_amg_sweep: out must not alias z or r");
  hipStream_t s = as_stream(stream);
  MG_DTYPE_DISPATCH(dtype, AMG_LANES_DISPATCH(lanes, ({
    const T *ae = (const T *)Ae, *av = (const T *)Av, *w = (const T *)wd, *rr = (const T *)r;
    if (mode == 0)
      k_amg_sweep<T, L, 0><<<amg_grid(n, L), kMB, 0, s>>>(n, rowptr, col, ae, av, w, rr,
                                                         (T *)z, (T *)out);
    else if (mode == 1)
      k_amg_sweep<T, L, 1><<<amg_grid(n, L), kMB, 0, s>>>(n, rowptr, col, ae, av, w, rr,
                                                         (T *)z, (T *)out);
    else
      k_amg_sweep<T, L, 2><<<amg_grid(n, L), kMB, 0, s>>>(n, rowptr, col, ae, av, w, rr,
                                                         (T *)z, (T *)out);
  })));
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

ESR_API int esr_amg_jacobi_post_sell(int dtype, int64_t n, const int64_t *soff,
                                     const uint8_t *cnt, const void *colv, int32_t idx16,
                                     const void *val, const void *Av, const void *wd,
                                     const void *r, const void *z, void *out, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1, "esr_amg_jacobi_post_sell: n must be >= 1");
  ESR_REQUIRE(soff && cnt && colv && val && Av && wd && r && z && out,
              "esr_amg_jacobi_post_sell: null pointer");
  ESR_REQUIRE(out != z, "esr_amg_jacobi_post_sell: out must not alias z");
  const int32_t *rowptr = nullptr, *col = nullptr;
  const void *Ae = nullptr;
  MG_DTYPE_DISPATCH(dtype, MG_OP_DISPATCH(T, (mg_post<T>(n, nullptr, A, wd, r, z, nullptr, out,
                                                         as_stream(stream)))));
  ESR_GUARD_END
}

ESR_API int esr_mg_dense(int dtype, int64_t n, const double *Ainv, const void *r, void *z,
                         void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(n >= 1, "esr_mg_dense: n must be >= 1");
  ESR_REQUIRE(Ainv && r && z && r != z, "esr_mg_dense: null or aliased pointer");
  MG_DTYPE_DISPATCH(dtype, (k_mg_dense<T><<<(unsigned)ceil_div(n * 64, kMB), kMB, 0,
                                            as_stream(stream)>>>(n, Ainv, (const T *)r, (T *)z)));
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}
