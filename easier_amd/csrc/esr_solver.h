// Device-resident state of the fused Krylov solvers (host and device view).
#pragma once
#include <cstdint>

namespace esr {

// Jacobi/identity-preconditioned CG (easier/numeric/solver/cg.py:74-137).
// Scalars live on the device so an iteration needs no host round trip; the
// host polls `done` once per captured chunk of iterations.
// Fields written by one workgroup of a kernel are never read by the other
// workgroups of that kernel: a kernel decides into the *_next fields and the
// next kernel commits them into the *_cur fields.
struct PcgState {
  double rz_cur;   // sum(r*z) of the current r, z (Step's rzsum)
  double rz_next;  // sum(r*z) of the step being run (committed by its second kernel)
  double pAp;
  double alpha;
  double rr;
  double rz_new;
  double rnorm;    // ||r|| after the last Step (the recursive residual)
  double bnorm;
  double tol;      // max(rtol*||b||, atol) if atol else rtol*||b||
  double rtol;
  double atol;     // <= 0: unset
  double beta;
  int64_t k_cur;   // `iters` of the reference loop at the current Step
  int64_t k_next;
  int64_t maxiter; // < 0: unset
  int64_t done;    // 1 once the loop has exited
  int64_t iters;   // value of `iters` at exit
  int64_t fresh;   // 1: the next step starts from p as it is (after Init or a re-arm)
  int64_t pbuf;    // which of the two direction buffers holds the current p
  int64_t pbuf_next;
  int64_t pad0;
};

// Restarted GMRES(m) Arnoldi cycle (easier/numeric/solver/gmres.py:214-266).
struct GmresState {
  double rnorm;       // true residual norm of the cycle's start
  double bnorm;
  double H[33 * 32];  // (m+1) x m, row-major with row stride 32 (m <= 32)
  int64_t broken;     // Arnoldi breakdown (h < 1e-15) seen at step jend
  int64_t jend;       // last Arnoldi step of the cycle
  // restart loop on the device (esr_gmres_loop_*): exit tolerance and cap,
  // cycles run, exit flag, and the cycle's least-squares solution
  double tol;
  int64_t maxiter;    // < 0: none
  int64_t iters;
  int64_t done;
  double y[32];
};

constexpr int kGmresMaxRestart = 32;

} // namespace esr
