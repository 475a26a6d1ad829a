// P1 (piecewise-linear) finite elements for -div(kappa grad u) = f on a
// triangle mesh with zero Dirichlet values on the boundary, assembled into the
// Linsys layout (easier/numeric/linsys.py:10-24): one directed edge per ordered
// pair of vertices that share a triangle edge, sorted by (row, col), with the
// off-diagonal stiffness Ae per edge and the diagonal Av per vertex.
//
// Element stiffness with kappa constant on triangle T (area |T|, edge vectors
// e_a = c - b, e_b = a - c, e_c = b - a opposite vertices a, b, c):
//   K_ij = kappa_T (e_i . e_j) / (4 |T|)
// Load: b = M f with the consistent mass matrix, M_ij = |T|/12 (1 + [i == j]).
// Contributions are added in increasing triangle order. Boundary vertices
// (endpoints of triangle edges that belong to one triangle) get identity rows
// (Av = 1, Ae = 0, b = 0) and their columns are zeroed in interior rows, which
// keeps the matrix symmetric positive definite.
#include <algorithm>
#include <cmath>
#include <vector>

#include "esr_host.h"

namespace {

// The distinct undirected triangle edges (a < b) in (a, b) order, with the
// number of triangles sharing each: a counting sort by a (one bucket per
// vertex), then each bucket -- a vertex's few larger neighbours -- sorted.
bool build_edges(int64_t np_, int64_t nc, const int64_t *cells, std::vector<int64_t> &ua,
                 std::vector<int64_t> &ub, std::vector<int64_t> &count) {
  std::vector<int64_t> head((size_t)np_ + 1, 0);
  for (int64_t t = 0; t < nc; ++t) {
    for (int k = 0; k < 3; ++k) {
      const int64_t u = cells[3 * t + k], v = cells[3 * t + (k + 1) % 3];
      if (u < 0 || u >= np_ || v < 0 || v >= np_ || u == v) return false;
      head[(size_t)std::min(u, v) + 1]++;
    }
  }
  for (int64_t a = 0; a < np_; ++a) head[a + 1] += head[a];
  std::vector<int64_t> pos(head.begin(), head.end() - 1), nb((size_t)head[np_]);
  for (int64_t t = 0; t < nc; ++t) {
    for (int k = 0; k < 3; ++k) {
      const int64_t u = cells[3 * t + k], v = cells[3 * t + (k + 1) % 3];
      nb[(size_t)pos[std::min(u, v)]++] = std::max(u, v);
    }
  }
  for (int64_t a = 0; a < np_; ++a) {
    int64_t *lo = nb.data() + head[a], *hi = nb.data() + head[a + 1];
    std::sort(lo, hi);
    for (int64_t *i = lo; i < hi;) {
      int64_t *j = i;
      while (j < hi && *j == *i) ++j;
      ua.push_back(a);
      ub.push_back(*i);
      count.push_back((int64_t)(j - i));
      i = j;
    }
  }
  return true;
}

} // namespace

ESR_HAPI int esr_p1_nnz(int64_t np_, int64_t nc, const int64_t *cells, int64_t *nnz) {
  ESR_HGUARD_BEGIN
  ESR_HREQUIRE(np_ >= 1 && nc >= 1 && cells && nnz, "esr_p1_nnz: bad arguments");
  std::vector<int64_t> ua, ub, cnt;
  ESR_HREQUIRE(build_edges(np_, nc, cells, ua, ub, cnt), "esr_p1_nnz: invalid cell indices");
  *nnz = 2 * (int64_t)ua.size();
  return ESR_OK;
  ESR_HGUARD_END
}

ESR_HAPI int esr_p1_assemble(int64_t np_, int64_t nc, const int64_t *cells,
                             const double *points, const double *kappa, const double *f,
                             int64_t nnz, int64_t *row, int64_t *col, double *Ae,
                             double *Av, double *b, uint8_t *boundary) {
  ESR_HGUARD_BEGIN
  ESR_HREQUIRE(np_ >= 1 && nc >= 1 && cells && points && row && col && Ae && Av,
               "esr_p1_assemble: null argument");
  std::vector<int64_t> ua, ub, cnt;
  ESR_HREQUIRE(build_edges(np_, nc, cells, ua, ub, cnt), "esr_p1_assemble: invalid cells");
  ESR_HREQUIRE(nnz == 2 * (int64_t)ua.size(), "esr_p1_assemble: nnz %lld != %lld",
               (long long)nnz, (long long)(2 * ua.size()));
  const int64_t nu = (int64_t)ua.size();
  // boundary vertices
  std::vector<uint8_t> bnd(np_, 0);
  for (int64_t k = 0; k < nu; ++k)
    if (cnt[k] == 1) bnd[ua[k]] = bnd[ub[k]] = 1;
  // CSR rows: neighbours ascending
  std::vector<int64_t> rowptr(np_ + 1, 0);
  for (int64_t k = 0; k < nu; ++k) {
    rowptr[ua[k] + 1]++;
    rowptr[ub[k] + 1]++;
  }
  for (int64_t v = 0; v < np_; ++v) rowptr[v + 1] += rowptr[v];
  std::vector<int64_t> cur(rowptr.begin(), rowptr.end() - 1);
  for (int64_t k = 0; k < nu; ++k) {
    col[cur[ua[k]]++] = ub[k];
    col[cur[ub[k]]++] = ua[k];
  }
  for (int64_t v = 0; v < np_; ++v) {
    std::sort(col + rowptr[v], col + rowptr[v + 1]);
    for (int64_t j = rowptr[v]; j < rowptr[v + 1]; ++j) row[j] = v;
  }
  auto slot = [&](int64_t r, int64_t c) -> int64_t {
    const int64_t *lo = col + rowptr[r], *hi = col + rowptr[r + 1];
    const int64_t *it = std::lower_bound(lo, hi, c);
    return it - col;
  };
  std::fill(Ae, Ae + nnz, 0.0);
  std::fill(Av, Av + np_, 0.0);
  if (b) std::fill(b, b + np_, 0.0);
  for (int64_t t = 0; t < nc; ++t) {
    const int64_t v[3] = {cells[3 * t], cells[3 * t + 1], cells[3 * t + 2]};
    const double xa = points[2 * v[0]], ya = points[2 * v[0] + 1];
    const double xb = points[2 * v[1]], yb = points[2 * v[1] + 1];
    const double xc = points[2 * v[2]], yc = points[2 * v[2] + 1];
    const double cross = (xb - xa) * (yc - ya) - (xc - xa) * (yb - ya);
    const double area = 0.5 * std::fabs(cross);
    const double kap = kappa ? kappa[t] : 1.0;
    const double s = kap / (4.0 * area);
    const double ex[3] = {xc - xb, xa - xc, xb - xa};
    const double ey[3] = {yc - yb, ya - yc, yb - ya};
    for (int i = 0; i < 3; ++i) {
      Av[v[i]] += s * (ex[i] * ex[i] + ey[i] * ey[i]);
      for (int j = 0; j < 3; ++j) {
        if (i == j) continue;
        Ae[slot(v[i], v[j])] += s * (ex[i] * ex[j] + ey[i] * ey[j]);
      }
    }
    if (b && f) {
      const double m = area / 12.0;
      for (int i = 0; i < 3; ++i)
        b[v[i]] += m * (2.0 * f[v[i]] + f[v[(i + 1) % 3]] + f[v[(i + 2) % 3]]);
    }
  }
  for (int64_t r = 0; r < np_; ++r) {
    if (bnd[r]) {
      Av[r] = 1.0;
      if (b) b[r] = 0.0;
    }
    for (int64_t j = rowptr[r]; j < rowptr[r + 1]; ++j)
      if (bnd[r] || bnd[col[j]]) Ae[j] = 0.0;
  }
  if (boundary) std::copy(bnd.begin(), bnd.end(), boundary);
  return ESR_OK;
  ESR_HGUARD_END
}
