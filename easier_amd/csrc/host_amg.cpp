// Smoothed-aggregation coarsening (Vanek, Mandel and Brezina, Computing 56,
// 1996) of a symmetric positive definite operator in the Linsys layout
// (easier/numeric/linsys.py:10-24: off-diagonal CSR Ae + diagonal Av), for
// the algebraic multigrid hierarchy of the multilevel preconditioner. The
// reference has no preconditioner of its own (its solvers take any callable
// M, easier/numeric/solver/cg.py:86); this is the set-up of one level, run on
// the host once per operator.
//
//   strength:   j ~ i  when a_ij != 0 and a_ij^2 >= theta^2 |a_ii a_jj|
//   aggregates: three passes in vertex order -- (1) a vertex whose strong
//               neighbours are all free starts an aggregate with them; (2) a
//               free vertex joins the pass-1 aggregate of its strongest
//               aggregated neighbour; (3) a vertex still free starts an
//               aggregate with its free strong neighbours. A vertex without
//               strong neighbours stays out (its row of P_tent is zero):
//               the smoother alone handles it (e.g. a Dirichlet identity row).
//   prolongator P = (I - omega D^-1 A) P_tent, P_tent[i, agg(i)] = 1,
//               omega = 4 / (3 rho), rho = max_i (|a_ii| + sum_j |a_ij|) / |a_ii|
//               (Gershgorin bound of D^-1 A).
//   coarse operator Ac = P^T A P, row I summed over i in P^T row I (ascending
//               i), j over {i} then row i's entries in CSR order, K over P row j:
//               acc[K] += (p_iI a_ij) p_jK. Each row is summed in that fixed
//               order whatever the thread count, so the result is deterministic.
//
// The plan holds the level's aggregates, P (CSR by fine row, columns
// ascending), R = P^T (CSR by coarse row, fine indices ascending) and Ac in
// the Linsys layout (off-diagonal columns ascending, diagonal apart).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <exception>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "esr_host.h"

namespace {

struct AmgPlan {
  int64_t n = 0, nc = 0;
  double rho = 0.0;
  std::vector<int32_t> agg;
  std::vector<int32_t> prow, pcol;
  std::vector<double> pval;
  std::vector<int32_t> rrow, rcol;
  std::vector<double> rval;
  std::vector<int32_t> crow, ccol;
  std::vector<double> cAe, cAv;
};

// Rows [0, nrows) in chunks handed to nthreads workers; chunk c's output is
// produced by one worker into out[c], so concatenating the chunks in order
// gives the same rows whatever the thread count.
struct RowChunks {
  std::vector<int32_t> len;      // entries per row
  std::vector<int32_t> col;
  std::vector<double> val;
};

template <typename RowFn>
void run_rows(int64_t nrows, int nthreads, std::vector<RowChunks> &out, RowFn make_worker) {
  const int64_t chunk = 4096;
  const int64_t nchunks = (nrows + chunk - 1) / chunk;
  out.assign((size_t)nchunks, RowChunks());
  std::atomic<int64_t> next(0);
  std::atomic<bool> failed(false);
  std::exception_ptr error;
  std::mutex error_lock;
  // a worker's exception (e.g. std::bad_alloc) is kept and rethrown on the
  // calling thread after every worker has joined
  auto work = [&]() {
    try {
      auto row = make_worker();   // per-worker scratch (sparse accumulator)
      for (;;) {
        const int64_t c = next.fetch_add(1);
        if (c >= nchunks || failed.load()) break;
        RowChunks &o = out[(size_t)c];
        const int64_t lo = c * chunk, hi = std::min(nrows, lo + chunk);
        o.len.reserve((size_t)(hi - lo));
        for (int64_t r = lo; r < hi; ++r) row(r, o);
      }
    } catch (...) {
      std::lock_guard<std::mutex> g(error_lock);
      if (!error) error = std::current_exception();
      failed.store(true);
    }
  };
  const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(nthreads, nchunks));
  if (nt == 1) {
    work();
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) th.emplace_back(work);
    for (auto &t : th) t.join();
  }
  if (error) std::rethrow_exception(error);
}

// Concatenate chunked rows into a CSR (rowptr of nrows + 1). False when the
// entries do not fit int32 row pointers.
bool gather_rows(const std::vector<RowChunks> &parts, int64_t nrows, std::vector<int32_t> &rowptr,
                 std::vector<int32_t> &col, std::vector<double> &val) {
  int64_t tot = 0;
  for (const auto &p : parts) tot += (int64_t)p.col.size();
  if (tot > (int64_t)INT32_MAX) return false;
  rowptr.assign((size_t)nrows + 1, 0);
  int64_t r = 0;
  tot = 0;
  for (const auto &p : parts) {
    for (int32_t l : p.len) {
      tot += l;
      rowptr[(size_t)++r] = (int32_t)tot;
    }
  }
  col.clear();
  val.clear();
  col.reserve((size_t)tot);
  val.reserve((size_t)tot);
  for (const auto &p : parts) {
    col.insert(col.end(), p.col.begin(), p.col.end());
    val.insert(val.end(), p.val.begin(), p.val.end());
  }
  return true;
}

// Sparse accumulator over [0, m): values summed in arrival order, touched
// columns emitted ascending.
struct Spa {
  std::vector<double> acc;
  std::vector<int32_t> mark;     // stamp of the row that touched it last
  std::vector<int32_t> touched;
  int32_t stamp = 0;
  explicit Spa(int64_t m) : acc((size_t)m, 0.0), mark((size_t)m, 0) {}
  void begin() {
    ++stamp;
    touched.clear();
  }
  void add(int32_t k, double v) {
    if (mark[(size_t)k] != stamp) {
      mark[(size_t)k] = stamp;
      acc[(size_t)k] = v;
      touched.push_back(k);
    } else {
      acc[(size_t)k] += v;
    }
  }
  void sorted() { std::sort(touched.begin(), touched.end()); }
};

int coarsen(AmgPlan &pl, int64_t n, const int32_t *rowptr, const int32_t *col, const double *Ae,
            const double *Av, double theta, int nthreads) {
  pl.n = n;
  // Gershgorin bound of D^-1 A and the prolongator's smoothing weight
  double rho = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    double s = std::fabs(Av[i]);
    for (int32_t e = rowptr[i]; e < rowptr[i + 1]; ++e) s += std::fabs(Ae[e]);
    rho = std::max(rho, s / std::fabs(Av[i]));
  }
  pl.rho = rho;
  const double omega = 4.0 / (3.0 * rho);
  // strength of connection
  const double th2 = theta * theta;
  auto strong = [&](int64_t i, int32_t e) {
    const double a = Ae[e];
    return a != 0.0 && a * a >= th2 * std::fabs(Av[i] * Av[col[e]]) && col[e] != i;
  };
  // aggregates: agg = -1 free; pass-2 joins kept apart (joined) so that a
  // vertex joins only a pass-1 aggregate
  std::vector<int32_t> &agg = pl.agg;
  agg.assign((size_t)n, -1);
  std::vector<uint8_t> has_strong((size_t)n, 0);
  int32_t na = 0;
  for (int64_t i = 0; i < n; ++i) {
    bool any = false, free_all = true;
    for (int32_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
      if (!strong(i, e)) continue;
      any = true;
      if (agg[(size_t)col[e]] >= 0) free_all = false;
    }
    has_strong[(size_t)i] = any;
    if (!any || agg[(size_t)i] >= 0 || !free_all) continue;
    agg[(size_t)i] = na;
    for (int32_t e = rowptr[i]; e < rowptr[i + 1]; ++e)
      if (strong(i, e)) agg[(size_t)col[e]] = na;
    ++na;
  }
  std::vector<int32_t> joined((size_t)n, -1);
  for (int64_t i = 0; i < n; ++i) {
    if (agg[(size_t)i] >= 0 || !has_strong[(size_t)i]) continue;
    double best = -1.0;
    for (int32_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
      if (!strong(i, e) || agg[(size_t)col[e]] < 0) continue;
      const double s = std::fabs(Ae[e]) / std::sqrt(std::fabs(Av[i] * Av[col[e]]));
      if (s > best) {
        best = s;
        joined[(size_t)i] = agg[(size_t)col[e]];
      }
    }
  }
  for (int64_t i = 0; i < n; ++i)
    if (joined[(size_t)i] >= 0) agg[(size_t)i] = joined[(size_t)i];
  for (int64_t i = 0; i < n; ++i) {
    if (agg[(size_t)i] >= 0 || !has_strong[(size_t)i]) continue;
    agg[(size_t)i] = na;
    for (int32_t e = rowptr[i]; e < rowptr[i + 1]; ++e)
      if (strong(i, e) && agg[(size_t)col[e]] < 0) agg[(size_t)col[e]] = na;
    ++na;
  }
  pl.nc = na;
  if (na == 0) return ESR_OK;
  // P = (I - omega D^-1 A) P_tent, row by row
  std::vector<RowChunks> parts;
  run_rows(n, nthreads, parts, [&]() {
    auto spa = std::make_shared<Spa>(na);
    return [&, spa](int64_t i, RowChunks &o) {
      Spa &s = *spa;
      s.begin();
      const double wi = omega / Av[i];
      const int32_t ai = agg[(size_t)i];
      if (ai >= 0) s.add(ai, 1.0 - wi * Av[i]);
      for (int32_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
        const int32_t aj = agg[(size_t)col[e]];
        if (aj >= 0 && Ae[e] != 0.0) s.add(aj, -wi * Ae[e]);
      }
      s.sorted();
      o.len.push_back((int32_t)s.touched.size());
      for (int32_t k : s.touched) {
        o.col.push_back(k);
        o.val.push_back(s.acc[(size_t)k]);
      }
    };
  });
  if (!gather_rows(parts, n, pl.prow, pl.pcol, pl.pval)) {
    esr::set_error("esr_amg_coarsen: the prolongator has more than 2^31 - 1 entries");
    return ESR_ERR_UNSUPPORTED;
  }
  // R = P^T by a counting sort over fine rows (fine indices ascending per row)
  pl.rrow.assign((size_t)na + 1, 0);
  for (int32_t k : pl.pcol) pl.rrow[(size_t)k + 1]++;
  for (int32_t I = 0; I < na; ++I) pl.rrow[(size_t)I + 1] += pl.rrow[(size_t)I];
  pl.rcol.resize(pl.pcol.size());
  pl.rval.resize(pl.pcol.size());
  {
    std::vector<int32_t> cur(pl.rrow.begin(), pl.rrow.end() - 1);
    for (int64_t i = 0; i < n; ++i) {
      for (int32_t e = pl.prow[(size_t)i]; e < pl.prow[(size_t)i + 1]; ++e) {
        const int32_t q = cur[(size_t)pl.pcol[(size_t)e]]++;
        pl.rcol[(size_t)q] = (int32_t)i;
        pl.rval[(size_t)q] = pl.pval[(size_t)e];
      }
    }
  }
  // Ac = P^T A P, one coarse row at a time; the diagonal kept apart
  std::vector<RowChunks> cparts;
  std::vector<double> diag((size_t)na, 0.0);
  const std::vector<int32_t> &prow = pl.prow, &pcol = pl.pcol, &rrow = pl.rrow, &rcol = pl.rcol;
  const std::vector<double> &pval = pl.pval, &rval = pl.rval;
  run_rows(na, nthreads, cparts, [&]() {
    auto spa = std::make_shared<Spa>(na);
    return [&, spa](int64_t I, RowChunks &o) {
      Spa &s = *spa;
      s.begin();
      for (int32_t q = rrow[(size_t)I]; q < rrow[(size_t)I + 1]; ++q) {
        const int32_t i = rcol[(size_t)q];
        const double piI = rval[(size_t)q];
        {
          const double w = piI * Av[i];
          for (int32_t f = prow[(size_t)i]; f < prow[(size_t)i + 1]; ++f)
            s.add(pcol[(size_t)f], w * pval[(size_t)f]);
        }
        for (int32_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
          if (Ae[e] == 0.0) continue;
          const int32_t j = col[e];
          const double w = piI * Ae[e];
          for (int32_t f = prow[(size_t)j]; f < prow[(size_t)j + 1]; ++f)
            s.add(pcol[(size_t)f], w * pval[(size_t)f]);
        }
      }
      s.sorted();
      int32_t len = 0;
      for (int32_t k : s.touched) {
        if (k == (int32_t)I) {
          diag[(size_t)I] = s.acc[(size_t)k];
          continue;
        }
        o.col.push_back(k);
        o.val.push_back(s.acc[(size_t)k]);
        ++len;
      }
      o.len.push_back(len);
    };
  });
  if (!gather_rows(cparts, na, pl.crow, pl.ccol, pl.cAe)) {
    esr::set_error("esr_amg_coarsen: the coarse operator has more than 2^31 - 1 entries");
    return ESR_ERR_UNSUPPORTED;
  }
  pl.cAv.swap(diag);
  for (int64_t I = 0; I < na; ++I) {
    if (!(pl.cAv[(size_t)I] > 0.0)) {
      esr::set_error("esr_amg_coarsen: coarse diagonal %lld is %g (the operator is not "
                     "positive definite)", (long long)I, pl.cAv[(size_t)I]);
      return ESR_ERR_INVALID;
    }
  }
  return ESR_OK;
}

} // namespace

ESR_HAPI int esr_amg_coarsen(int64_t n, const int32_t *rowptr, const int32_t *col,
                             const double *Ae, const double *Av, double theta, int32_t nthreads,
                             void **plan, int64_t *sizes, double *rho) {
  ESR_HGUARD_BEGIN
  ESR_HREQUIRE(n >= 1 && n < (int64_t)INT32_MAX, "esr_amg_coarsen: n %lld out of range",
               (long long)n);
  ESR_HREQUIRE(rowptr && Av && plan && sizes && rho, "esr_amg_coarsen: null argument");
  ESR_HREQUIRE(theta >= 0.0 && theta < 1.0, "esr_amg_coarsen: theta %g outside [0, 1)", theta);
  *plan = nullptr;
  ESR_HREQUIRE(rowptr[0] == 0, "esr_amg_coarsen: rowptr[0] = %d", (int)rowptr[0]);
  for (int64_t i = 0; i < n; ++i) {
    ESR_HREQUIRE(rowptr[i + 1] >= rowptr[i], "esr_amg_coarsen: rowptr decreases at row %lld",
                 (long long)i);
    ESR_HREQUIRE(Av[i] > 0.0, "esr_amg_coarsen: diagonal %lld is %g (not positive)",
                 (long long)i, Av[i]);
  }
  ESR_HREQUIRE(rowptr[n] == 0 || (col && Ae), "esr_amg_coarsen: null col or Ae");
  for (int32_t e = 0; e < rowptr[n]; ++e)
    ESR_HREQUIRE(col[e] >= 0 && col[e] < n, "esr_amg_coarsen: column %d out of range",
                 (int)col[e]);
  auto *pl = new AmgPlan();
  const int rc = coarsen(*pl, n, rowptr, col, Ae, Av, theta, nthreads > 0 ? nthreads : 1);
  if (rc != ESR_OK) {
    delete pl;
    return rc;
  }
  sizes[0] = pl->nc;
  sizes[1] = (int64_t)pl->pcol.size();
  sizes[2] = (int64_t)pl->ccol.size();
  *rho = pl->rho;
  *plan = pl;
  return ESR_OK;
  ESR_HGUARD_END
}

ESR_HAPI int esr_amg_fetch(void *plan, int32_t *agg, int32_t *p_rowptr, int32_t *p_col,
                           double *p_val, int32_t *r_rowptr, int32_t *r_col, double *r_val,
                           int32_t *c_rowptr, int32_t *c_col, double *c_Ae, double *c_Av) {
  ESR_HGUARD_BEGIN
  ESR_HREQUIRE(plan, "esr_amg_fetch: null plan");
  const AmgPlan &pl = *static_cast<const AmgPlan *>(plan);
  auto put = [](const auto &v, auto *dst) {
    if (dst) std::copy(v.begin(), v.end(), dst);
  };
  put(pl.agg, agg);
  if (pl.nc > 0) {
    put(pl.prow, p_rowptr);
    put(pl.pcol, p_col);
    put(pl.pval, p_val);
    put(pl.rrow, r_rowptr);
    put(pl.rcol, r_col);
    put(pl.rval, r_val);
    put(pl.crow, c_rowptr);
    put(pl.ccol, c_col);
    put(pl.cAe, c_Ae);
    put(pl.cAv, c_Av);
  }
  return ESR_OK;
  ESR_HGUARD_END
}

ESR_HAPI int esr_amg_destroy(void *plan) {
  delete static_cast<AmgPlan *>(plan);
  return ESR_OK;
}
