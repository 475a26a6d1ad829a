// Drives every host entry point of libesrhip.so (mesh generators, stable CSR
// build, RCM, block order, bandwidth, P1 assembly, smoothed-aggregation
// levels) over small and odd sizes,
// including the error paths, for a build with AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/test_sanitize.py). Exit 0 when every
// call behaved; the sanitizers abort on the first memory or UB error.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <vector>

#include "esr_hip.h"

namespace esr {
// the library's error sink (defined in its runtime for the real build)
void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  char buf[512];
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
}
} // namespace esr

#define CHECK(c)                                                               \
  do {                                                                         \
    if (!(c)) {                                                                \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__);     \
      return 1;                                                                \
    }                                                                          \
  } while (0)

static int mesh_case(int kind, int64_t n, double jitter) {
  int64_t np_, nc, ne, nbc;
  CHECK(esr_mesh_sizes(kind, n, &np_, &nc, &ne, &nbc) == ESR_OK);
  std::vector<int64_t> src(ne), dst(ne), cells(nc * 3), bcells(nbc), bpoints(nbc * 2);
  std::vector<double> points(np_ * 2);
  if (kind == 4)
    CHECK(esr_mesh_tri4(n, src.data(), dst.data(), cells.data(), bcells.data(),
                        bpoints.data(), points.data()) == ESR_OK);
  else
    CHECK(esr_mesh_tri2(n, jitter, 7, src.data(), dst.data(), cells.data(), bcells.data(),
                        bpoints.data(), points.data()) == ESR_OK);
  // stable CSR of the edges by destination
  std::vector<int64_t> rowptr(nc + 1), perm(ne);
  CHECK(esr_csr_from_idx(dst.data(), ne, nc, rowptr.data(), perm.data()) == ESR_OK);
  CHECK(rowptr[nc] == ne);
  // RCM and blocks over the dual graph
  std::vector<int64_t> order(nc);
  CHECK(esr_rcm(nc, src.data(), dst.data(), ne, order.data()) == ESR_OK);
  std::vector<int64_t> seen(nc, 0);
  for (int64_t v : order) CHECK(v >= 0 && v < nc && !seen[v]++);
  for (int64_t blk : {1, 3, 64, 256}) {
    CHECK(esr_block_order(nc, src.data(), dst.data(), ne, blk, order.data()) == ESR_OK);
    std::fill(seen.begin(), seen.end(), 0);
    for (int64_t v : order) CHECK(v >= 0 && v < nc && !seen[v]++);
  }
  int64_t bw = -1;
  CHECK(esr_bandwidth(src.data(), dst.data(), ne, &bw) == ESR_OK && bw >= 0);
  // P1 assembly with and without a coefficient and a right-hand side
  int64_t nnz = 0;
  CHECK(esr_p1_nnz(np_, nc, cells.data(), &nnz) == ESR_OK);
  std::vector<int64_t> row(nnz), col(nnz);
  std::vector<double> Ae(nnz), Av(np_), b(np_), f(np_, 0.5), kappa(nc, 2.0);
  std::vector<uint8_t> bnd(np_);
  CHECK(esr_p1_assemble(np_, nc, cells.data(), points.data(), kappa.data(), f.data(), nnz,
                        row.data(), col.data(), Ae.data(), Av.data(), b.data(),
                        bnd.data()) == ESR_OK);
  CHECK(esr_p1_assemble(np_, nc, cells.data(), points.data(), nullptr, nullptr, nnz,
                        row.data(), col.data(), Ae.data(), Av.data(), nullptr,
                        bnd.data()) == ESR_OK);
  // smoothed-aggregation levels down to a few rows, on 1 and 3 host threads
  std::vector<int32_t> rp(np_ + 1, 0), cl(nnz);
  for (int64_t e = 0; e < nnz; ++e) {
    rp[row[e] + 1]++;
    cl[e] = (int32_t)col[e];
  }
  for (int64_t v = 0; v < np_; ++v) rp[v + 1] += rp[v];
  std::vector<double> ae(Ae), av(Av);
  for (int level = 0; level < 8; ++level) {
    const int64_t m = (int64_t)av.size();
    std::vector<int32_t> agg(m);
    for (int nt : {1, 3}) {
      void *plan = nullptr;
      int64_t sz[3];
      double rho = 0.0;
      CHECK(esr_amg_coarsen(m, rp.data(), cl.data(), ae.data(), av.data(), 0.08, nt, &plan, sz,
                            &rho) == ESR_OK);
      CHECK(plan && rho >= 1.0);
      const int64_t c = sz[0];
      std::vector<int32_t> prow(m + 1), pcol(sz[1]), rrow(c + 1), rcol(sz[1]), crow(c + 1),
          ccol(sz[2]);
      std::vector<double> pval(sz[1]), rval(sz[1]), cae(sz[2]), cav(c);
      CHECK(esr_amg_fetch(plan, agg.data(), prow.data(), pcol.data(), pval.data(), rrow.data(),
                          rcol.data(), rval.data(), crow.data(), ccol.data(), cae.data(),
                          cav.data()) == ESR_OK);
      CHECK(esr_amg_fetch(plan, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                          nullptr, nullptr, nullptr, nullptr) == ESR_OK);
      CHECK(esr_amg_destroy(plan) == ESR_OK);
      if (nt == 3) {
        if (c == 0 || c >= m) return 0;
        rp.swap(crow);
        cl.swap(ccol);
        ae.swap(cae);
        av.swap(cav);
      }
    }
  }
  return 0;
}

int main() {
  for (int64_t n : {1, 2, 5, 17})
    for (int kind : {2, 4})
      if (mesh_case(kind, n, 0.0)) return 1;
  if (mesh_case(2, 9, 0.2)) return 1;
  // error paths: out-of-range indices, bad sizes, wrong nnz
  std::vector<int64_t> idx = {0, 3, 1}, rowptr(3), perm(3), order(2);
  CHECK(esr_csr_from_idx(idx.data(), 3, 2, rowptr.data(), perm.data()) == ESR_ERR_INVALID);
  std::vector<int64_t> a = {0, 5}, b = {1, 0};
  CHECK(esr_rcm(2, a.data(), b.data(), 2, order.data()) == ESR_ERR_INVALID);
  CHECK(esr_block_order(2, a.data(), b.data(), 2, 4, order.data()) == ESR_ERR_INVALID);
  CHECK(esr_block_order(2, b.data(), b.data(), 2, 0, order.data()) == ESR_ERR_INVALID);
  int64_t s0, s1, s2, s3;
  CHECK(esr_mesh_sizes(3, 4, &s0, &s1, &s2, &s3) != ESR_OK);
  CHECK(esr_mesh_sizes(2, 0, &s0, &s1, &s2, &s3) != ESR_OK);
  // a zero diagonal, a column out of range and an indefinite operator
  std::vector<int32_t> rp2 = {0, 1, 2}, cl2 = {1, 0}, cl3 = {1, 2};
  std::vector<double> ae2 = {-1.0, -1.0}, av2 = {2.0, 0.0}, av3 = {1.0, 1.0},
                      ae3 = {-4.0, -4.0};
  void *plan = nullptr;
  int64_t sz[3];
  double rho;
  CHECK(esr_amg_coarsen(2, rp2.data(), cl2.data(), ae2.data(), av2.data(), 0.08, 1, &plan, sz,
                        &rho) == ESR_ERR_INVALID && plan == nullptr);
  CHECK(esr_amg_coarsen(2, rp2.data(), cl3.data(), ae2.data(), av3.data(), 0.08, 1, &plan, sz,
                        &rho) == ESR_ERR_INVALID && plan == nullptr);
  CHECK(esr_amg_coarsen(2, rp2.data(), cl2.data(), ae3.data(), av3.data(), 0.08, 1, &plan, sz,
                        &rho) == ESR_ERR_INVALID && plan == nullptr);
  CHECK(esr_amg_fetch(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                      nullptr, nullptr, nullptr, nullptr) == ESR_ERR_INVALID);
  std::printf("host entry points clean\n");
  return 0;
}
