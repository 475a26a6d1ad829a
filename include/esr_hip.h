/*
 * esr_hip.h -- C ABI of libesrhip.so, the MI355X (gfx950) backend of the
 * EASIER programming model (esr.Module / esr.Selector / esr.Reducer /
 * esr.compile, easier.numeric CG/GMRES).
 *
 * Contract for every entry point:
 *   - returns 0 on success, a positive ESR_ERR_* code otherwise;
 *     esr_last_error() then returns a thread-local message;
 *   - no C++ exception crosses this boundary;
 *   - all device buffers are owned by the caller (plain device pointers and
 *     sizes, no framework types); the library allocates device memory only
 *     inside objects it hands back (compiled modules, graph executables), each
 *     with an explicit destroy call, plus one fixed device array per GPU for
 *     esr_aggregate_noalloc (ESR_AGG_SCRATCH_BYTES, held until exit);
 *   - every device launch goes on the caller's stream (`stream` is a
 *     hipStream_t passed as void*; NULL means the null stream);
 *   - index arrays are int32 on the device; their ranges are validated by the
 *     host before any launch (kernels do not bounds-check indices).
 *
 * Each declaration cites the reference interface (file:line in
 * jack-pan-ai/EASIER) whose behaviour it replaces.
 */
#ifndef ESR_HIP_H
#define ESR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define ESR_OK 0
#define ESR_ERR_INVALID 1     /* bad argument (shape, dtype, size)          */
#define ESR_ERR_HIP 2         /* HIP runtime error                          */
#define ESR_ERR_UNSUPPORTED 3 /* valid request this build does not handle   */
#define ESR_ERR_RTC 4         /* runtime compilation failed (log in error)  */
#define ESR_ERR_INTERNAL 5    /* unexpected internal failure                */

/* ---- dtype codes --------------------------------------------------------- */
#define ESR_F32 0
#define ESR_F64 1
#define ESR_I32 2
#define ESR_I64 3
#define ESR_U8 4 /* torch.bool / uint8 */
#define ESR_BF16 5

/* ---- Reducer ops: easier/core/module.py:386 (ReduceOp) ------------------- */
#define ESR_RED_SUM 0
#define ESR_RED_PROD 1
#define ESR_RED_MEAN 2
#define ESR_RED_AMAX 3
#define ESR_RED_AMIN 4

/* ---- aggregators: easier/core/module.py:825-884 -------------------------- */
#define ESR_AGG_SUM 0
#define ESR_AGG_PROD 1
#define ESR_AGG_NORM 2 /* p-norm, p passed separately (inf allowed)        */
#define ESR_AGG_MAX 3
#define ESR_AGG_MIN 4

/* ---- library ------------------------------------------------------------- */
int esr_abi_version(void);
const char *esr_last_error(void);
/* gcnArchName of the current device, e.g. "gfx950:sramecc+:xnack-". */
int esr_device_arch(char *buf, int buflen);
/* Blocks until all work on `stream` is done. */
int esr_stream_sync(void *stream);

/* ---- a1: Selector.forward  (easier/core/module.py:364-371) ---------------
 * dst[e, :] = src[idx[e], :] for e < n_out; a row is `width` elements of
 * `dtype`. Pure data movement: bit-exact.                                    */
int esr_gather(int dtype, const void *src, int64_t n_src, int64_t width,
               const int32_t *idx, int64_t n_out, void *dst, void *stream);

/* ---- a2: Reducer.forward  (easier/core/module.py:441-461) -----------------
 * out[r, :] = reduce_{j in [rowptr[r], rowptr[r+1])} vals[perm ? perm[j] : j, :]
 * i.e. torch scatter_reduce_(0, idx, vals, op, include_self=False) on a CSR
 * view of idx (rows = distinct targets, entries in original edge order).
 * Rows of at most ESR_SEQ_ROW_MAX entries are reduced sequentially in entry
 * order from the op's identity (bit-identical to torch CPU); longer rows, listed
 * in long_rows[0..n_long), use a fixed-shape tree (deterministic, not
 * bit-identical). keep_empty != 0: rows with no entries keep `out` (the
 * `out=` argument); keep_empty == 0: they are set to 0.                      */
#define ESR_SEQ_ROW_MAX 256
int esr_csr_reduce(int op, int dtype, const void *vals, int64_t width,
                   const int32_t *rowptr, const int32_t *perm, int64_t n_rows,
                   const int32_t *long_rows, int64_t n_long, int keep_empty,
                   void *out, void *stream);

/* ---- a4: aggregators esr.sum/prod/norm/max/min (module.py:825-884, with the
 * single-partition rewrite of passes/dataflow_distribution.py:197-289) ------
 * out[0, c] = agg_{i<n} x[i, c] for c < width (x row-major (n, width)).
 * Two fixed-shape stages through `ws` (esr_aggregate_ws_bytes bytes):
 * deterministic run to run.                                                  */
int esr_aggregate_ws_bytes(int dtype, int64_t n, int64_t width, int64_t *bytes);
int esr_aggregate(int op, int dtype, const void *x, int64_t n, int64_t width,
                  double p, void *out, void *ws, int64_t ws_bytes,
                  void *stream);
/* esr_aggregate without a workspace argument and without any allocation.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). The stage-1 partials go to a device array of
 * ESR_AGG_SCRATCH_BYTES bytes that the library holds on every GPU it runs on
 * until the process ends; the device address is looked up on the first call
 * on each device. Requests whose workspace (esr_aggregate_ws_bytes) exceeds
 * that size are rejected with ESR_ERR_INVALID (f64: any n for width <= 16).
 * Every call on one GPU shares the array, so calls that overlap in time on
 * one GPU (two streams, two host threads) race; calls ordered on one stream
 * do not. Results are bit-identical to esr_aggregate's.                     */
#define ESR_AGG_SCRATCH_BYTES 131072
int esr_aggregate_noalloc(int op, int dtype, const void *x, int64_t n,
                          int64_t width, double p, void *out, void *stream);

/* ---- elementwise operators of a traced esr.Module forward ----------------
 * Executes one torch elementwise operator (the call_function/call_method
 * nodes evaluated by runtime/jit_engine/values.py:330-383 in the reference)
 * over an up-to-6-d strided iteration space.                                 */
#define ESR_EW_MAXDIM 6
/* unary */
#define ESR_EW_COPY 0 /* cast/copy                                          */
#define ESR_EW_NEG 1
#define ESR_EW_ABS 2
#define ESR_EW_EXP 3
#define ESR_EW_LOG 4
#define ESR_EW_SQRT 5
#define ESR_EW_RSQRT 6
#define ESR_EW_SIN 7
#define ESR_EW_COS 8
#define ESR_EW_TANH 9
#define ESR_EW_SIGN 10
#define ESR_EW_RECIP 11
#define ESR_EW_FLOOR 12
#define ESR_EW_CEIL 13
#define ESR_EW_NOT 14
#define ESR_EW_SIGMOID 15
#define ESR_EW_RELU 16
#define ESR_EW_ISNAN 17
/* binary */
#define ESR_EW_ADD 32
#define ESR_EW_SUB 33
#define ESR_EW_MUL 34
#define ESR_EW_DIV 35
#define ESR_EW_POW 36
#define ESR_EW_MAXIMUM 37
#define ESR_EW_MINIMUM 38
#define ESR_EW_LT 39
#define ESR_EW_LE 40
#define ESR_EW_GT 41
#define ESR_EW_GE 42
#define ESR_EW_EQ 43
#define ESR_EW_NE 44
#define ESR_EW_AND 45
#define ESR_EW_OR 46
#define ESR_EW_ATAN2 47
#define ESR_EW_FMOD 48
#define ESR_EW_FLOORDIV 49
/* ternary */
#define ESR_EW_WHERE 64 /* in0 (u8) ? in1 : in2                            */

/* out and every tensor input share `shape` (ndim <= 6) through their own
 * element strides (stride 0 = broadcast). Tensor inputs all have dtype
 * `in_dtype`; an input with a NULL pointer is the host scalar scalars[k]
 * (cast to in_dtype). For WHERE, input 0 is u8 and inputs 1, 2 are in_dtype.
 * out_dtype is in_dtype, or ESR_U8 for comparisons/ISNAN/NOT/AND/OR, or any
 * dtype for COPY (a cast).                                                   */
int esr_ewise(int op, int ndim, const int64_t *shape, int out_dtype, void *out,
              const int64_t *out_strides, int in_dtype, int n_in,
              void *const *in_ptrs, const int64_t *in_strides,
              const double *scalars, void *stream);
int esr_fill(int dtype, void *dst, int64_t n, double value, void *stream);

/* y[r] = sum_{k<kdim} A[r*lda + k*ldk] * x[k*incx] summed in k order, for
 * r < n: the column-block product torch.matmul(V[:, :i], y[:i]) of
 * easier/numeric/solver/gmres.py:133-136 (UpdateX).                          */
int esr_gemv_rows(int dtype, const void *A, int64_t n, int64_t kdim,
                  int64_t lda, int64_t ldk, const void *x, int64_t incx,
                  void *y, void *stream);

/* ---- runtime compilation and launch of fused kernels ---------------------
 * Replaces the reference's per-node interpreter (jit_engine.py:1114-1172) by
 * fused kernels compiled once per esr.compile() (hipRTC, gfx950).            */
int esr_rtc_compile(const char *source, const char *name, const char *options,
                    void **module_out);
int esr_rtc_check(const char *source, const char *options,
                  int64_t *code_bytes); /* compile only, no device needed */
int esr_rtc_function(void *module, const char *kernel_name, void **fn_out);
int esr_rtc_unload(void *module);
int esr_launch(void *fn, uint32_t gx, uint32_t gy, uint32_t gz, uint32_t bx,
               uint32_t by, uint32_t bz, uint32_t shmem, void *stream,
               void *argbuf, size_t argbytes);

/* ---- HIP graphs (capture of a fixed launch sequence) --------------------- */
int esr_graph_begin(void *stream);
int esr_graph_end(void *stream, void **exec_out);
int esr_graph_launch(void *exec, void *stream);
int esr_graph_destroy(void *exec);

/* ---- a9: mesh generators (host) -------------------------------------------
 * esr_mesh_tri4: csrc/triangular_mesh.cpp:10-142 (4 triangles per square
 * around its centre point; same arrays, same numbering, same edge order).
 * esr_mesh_tri2: the same six arrays for 2 triangles per square (diagonal
 * from (i,j) to (i+1,j+1)), optionally with interior points jittered.
 * Sizes: esr_mesh_sizes(kind, n, &np, &nc, &ne, &nbc), kind 4 or 2.
 * Arrays (caller-allocated, host memory): src/dst (ne) int64, cells (nc*3)
 * int64, bcells (nbc) int64, bpoints (nbc*2) int64, points (np*2) double.   */
int esr_mesh_sizes(int kind, int64_t n, int64_t *np_, int64_t *nc,
                   int64_t *ne, int64_t *nbc);
int esr_mesh_tri4(int64_t n, int64_t *src, int64_t *dst, int64_t *cells,
                  int64_t *bcells, int64_t *bpoints, double *points);
int esr_mesh_tri2(int64_t n, double jitter, uint64_t seed, int64_t *src,
                  int64_t *dst, int64_t *cells, int64_t *bcells,
                  int64_t *bpoints, double *points);

/* ---- a10: index renumbering (host) -----------------------------------------
 * Replaces the world-size-1 outcome of passes/sparse_encoding/
 * sparse_encoding.py:482-595 (edges stable-sorted by the bound Reducer's
 * idx) and adds a locality renumbering of the vertex/cell space.
 *
 * esr_csr_from_idx: rowptr[n+1] and perm[m] such that perm lists the entries
 * of idx grouped by row, rows ascending, entries of one row in input order
 * (a stable counting sort). Fails if any idx is outside [0, n).               */
int esr_csr_from_idx(const int64_t *idx, int64_t m, int64_t n, int64_t *rowptr,
                     int64_t *perm);
/* Reverse Cuthill-McKee order of the symmetric graph given by the directed
 * pair list (a[k], b[k]), k < m, on n vertices (self loops ignored; each
 * connected component started from a pseudo-peripheral vertex).
 * order[new] = old.                                                           */
int esr_rcm(int64_t n, const int64_t *a, const int64_t *b, int64_t m,
            int64_t *order);
/* Blocks for locality within one partition (the reference partitions
 * vertices among ranks, core/distpart.py:661-775, and keeps each rank's
 * order; on one GPU the unit of locality is a workgroup's tile of rows):
 * order[new] = old, consecutive runs of `block` vertices are breadth-first
 * balls grown over not yet placed vertices, started at the first unplaced
 * vertex of the RCM order (so blocks follow the RCM sweep), each ball in BFS
 * order. A ball that runs out of reachable vertices ends early; the balls
 * after it fill the rest of its run, so every run of `block` starts a ball. */
int esr_block_order(int64_t n, const int64_t *a, const int64_t *b, int64_t m,
                    int64_t block, int64_t *order);
/* Largest |row(e) - col(e)| over the pairs: the bandwidth RCM minimises.     */
int esr_bandwidth(const int64_t *a, const int64_t *b, int64_t m,
                  int64_t *bw);

/* ---- a12: P1 finite elements (host) ----------------------------------------
 * The Poisson operator in the Linsys layout the reference's Poisson tutorial
 * feeds to CG/GMRES (tutorial/poisson/poisson_main.py:28-44: Reducer(row),
 * Selector(col), Ae per directed edge, Av per vertex), here assembled from P1
 * elements (the reference assembles a cell-centred finite-volume operator,
 * tutorial/poisson/assemble_poisson.py:160-230). esr_p1_nnz gives the number
 * of directed vertex edges; esr_p1_assemble fills row/col (sorted by row,
 * then col), Ae, Av, b = M f (consistent mass matrix; NULL f or b skips it)
 * and the boundary mask, with zero Dirichlet rows/columns. kappa (per cell)
 * may be NULL (kappa = 1).                                                   */
int esr_p1_nnz(int64_t np_, int64_t nc, const int64_t *cells, int64_t *nnz);
int esr_p1_assemble(int64_t np_, int64_t nc, const int64_t *cells,
                    const double *points, const double *kappa, const double *f,
                    int64_t nnz, int64_t *row, int64_t *col, double *Ae,
                    double *Av, double *b, uint8_t *boundary);

/* ---- a3/a5: fused Jacobi/identity-preconditioned CG ------------------------
 * easier/numeric/solver/cg.py:12-137 on a Linsys (linsys.py:23-24) given as
 * a row-sorted CSR (rowptr, col = Selector index, Ae, Av). diag (nullable) is
 * the Jacobi diagonal: M(r) = r / diag. State lives on the device
 * (esr_pcg_workspace sizes it): esr_pcg_init runs Init and fixes
 * tol = max(rtol*||b||, atol) if atol > 0 else rtol*||b|| and maxiter (< 0:
 * none); esr_pcg_steps enqueues `count` iterations of Step + loop test +
 * Update (two kernels each, the Update of one step fused with the next
 * step's product on systems of at most 2^17 rows; the workspace holds the
 * second direction buffer) and one final loop test that also leaves the last
 * direction in p once the loop has exited, each a no-op once the loop has
 * exited (safe to capture in a graph);
 * esr_pcg_read copies scalars {rnorm, bnorm, rz, tol, alpha, beta} and ints
 * {iters, done, k} to the host. dtype: ESR_F64 or ESR_F32 (dots always
 * accumulate in double).                                                     */
int esr_pcg_workspace(int64_t n, int64_t *state_bytes, int64_t *ws_bytes);
int esr_pcg_init(int dtype, int64_t n, const int32_t *rowptr,
                 const int32_t *col, const void *Ae, const void *Av,
                 const void *diag, const void *b, void *x, void *r, void *z,
                 void *p, void *state, void *ws, double rtol, double atol,
                 int64_t maxiter, void *stream);
int esr_pcg_steps(int dtype, int64_t n, const int32_t *rowptr,
                  const int32_t *col, const void *Ae, const void *Av,
                  const void *diag, void *x, void *r, void *z, void *p,
                  void *Ap, void *state, void *ws, int64_t count,
                  void *stream);
/* Continue a stopped PCG after the caller replaced r by a freshly computed
 * residual (a reliable update of mixed-precision CG; no reference
 * counterpart): z = M r, beta = r.z / rz, p = z + beta p, new absolute exit
 * tolerance; the step counter continues.                                     */
int esr_pcg_rearm(int dtype, int64_t n, const void *diag, const void *r, void *z,
                  void *p, void *state, void *ws, double tol, void *stream);
int esr_pcg_read(const void *state, double *scalars, int64_t *ints,
                 void *stream);

/* ---- a3: the Linsys operator in sliced ELLPACK (SELL-64) ------------------
 * The same operator as the CSR above (linsys.py:23-24: the Reducer's rows,
 * the Selector's index, Ae), stored slice by slice: rows 64s..64s+63 form
 * slice s, its width w_s the longest of those rows, entry k of row 64s+l at
 * soff[s] + 64k + l; cnt[r] (uint8) the row's length; padded entries hold
 * value 0. Column indices are int16 offsets from the row when every one fits
 * (idx16 = 1), else int32 indices.
 * esr_sell_plan: soff (device, ceil(n/64)+1 int64) from rowptr/col; flags
 * (device, 2 int32) is scratch; returns the total slot count and idx16 on the
 * host (synchronises the stream); ESR_ERR_UNSUPPORTED for a row of more than
 * 255 entries. esr_sell_fill: cnt (n bytes), colv (slots int16 or int32),
 * val (slots values of dtype) from the CSR. esr_sell_values: val again from
 * new Ae values (same structure). esr_pcg_steps_sell: esr_pcg_steps with the
 * product on the SELL form (identical results). No reference counterpart:
 * the reference applies the operator as torch index_select + scatter_reduce
 * (core/module.py:364-371, 441-461).                                        */
int esr_sell_plan(int64_t n, const int32_t *rowptr, const int32_t *col,
                  int64_t *soff, int32_t *flags, int64_t *slots, int32_t *idx16,
                  void *stream);
int esr_sell_fill(int dtype, int64_t n, const int32_t *rowptr,
                  const int32_t *col, const void *Ae, const int64_t *soff,
                  int32_t idx16, uint8_t *cnt, void *colv, void *val,
                  void *stream);
int esr_sell_values(int dtype, int64_t n, const int32_t *rowptr,
                    const void *Ae, const int64_t *soff, void *val,
                    void *stream);
int esr_pcg_steps_sell(int dtype, int64_t n, const int64_t *soff,
                       const uint8_t *cnt, const void *colv, int32_t idx16,
                       const void *val, const void *Av, const void *diag,
                       void *x, void *r, void *z, void *p, void *Ap,
                       void *state, void *ws, int64_t count, void *stream);

/* ---- a6: fused GMRES(m) Arnoldi cycle ---------------------------------------
 * easier/numeric/solver/gmres.py:12-266 (m <= 32). esr_gmres_begin: ||b||;
 * esr_gmres_residual: r = b - A x, ||r|| (UpdateRnorm); esr_gmres_cycle:
 * V_0 = r/||r||, then for j < m: w = A(M(V_j)), orthogonalised against
 * V_0..V_j into H[:, j] (the reference's modified Gram-Schmidt, here classical
 * Gram-Schmidt applied twice: three reductions per step), H[j+1, j] = ||w||,
 * V_{j+1} = w / H[j+1, j], stopping
 * (later kernels no-op) at H[j+1, j] < 1e-15; V is m x n (one basis vector per
 * row), w0/w1 scratch n-vectors. esr_gmres_read: scalars {rnorm, bnorm}, ints
 * {jend, broken}, H ((m+1) x m row-major). esr_gmres_update_x:
 * x += M(sum_{i<jcount} V_i y_i) (UpdateX).                                  */
int esr_gmres_workspace(int64_t n, int64_t restart, int64_t *state_bytes,
                        int64_t *ws_bytes);
int esr_gmres_begin(int dtype, int64_t n, const void *b, void *state,
                    void *ws, void *stream);
int esr_gmres_residual(int dtype, int64_t n, const int32_t *rowptr,
                       const int32_t *col, const void *Ae, const void *Av,
                       const void *b, const void *x, void *r, void *state,
                       void *ws, void *stream);
int esr_gmres_cycle(int dtype, int64_t n, int64_t restart,
                    const int32_t *rowptr, const int32_t *col, const void *Ae,
                    const void *Av, const void *diag, const void *r, void *V,
                    void *w0, void *w1, void *state, void *ws, void *stream);
int esr_gmres_update_x(int dtype, int64_t n, int64_t jcount, const void *V,
                       const void *y, const void *diag, void *x, void *stream);
/* The restart loop of gmres.py:214-266 on the device (the host loop above
 * reads H and solves the least-squares problem on the host every cycle):
 * esr_gmres_loop_arm after esr_gmres_begin sets tol = max(rtol ||b||, atol)
 * (atol <= 0: rtol ||b||) and the cycle cap (maxiter < 0: none);
 * esr_gmres_loop_cycles enqueues `count` restart cycles, each the true
 * residual r = b - A x with the loop test, the Arnoldi cycle, the
 * least-squares solve min ||rnorm e_0 - H y|| (Givens rotations; the
 * reference uses a thin SVD, gmres.py:255-258) and x += M(V y), every kernel
 * a no-op once the loop has exited (safe to capture and replay);
 * esr_gmres_loop_read: scalars {rnorm, bnorm}, ints {iters, done}.        */
int esr_gmres_loop_arm(void *state, double rtol, double atol, int64_t maxiter,
                       void *stream);
int esr_gmres_loop_cycles(int dtype, int64_t n, int64_t restart,
                          const int32_t *rowptr, const int32_t *col,
                          const void *Ae, const void *Av, const void *diag,
                          const void *b, void *x, void *r, void *V, void *w0,
                          void *w1, void *state, void *ws, int64_t count,
                          void *stream);
int esr_gmres_loop_read(const void *state, double *scalars, int64_t *ints,
                        void *stream);
int esr_gmres_read(const void *state, int64_t restart, double *scalars,
                   int64_t *ints, double *H, void *stream);

/* ---- mixed-precision iterative refinement (new; no reference counterpart:
 * the reference runs every solver in the tensors' one dtype) ----------------
 * Outer loop in float64 around the float32 fused CG above:
 * esr_ir_residual: r = b - A x in float64 on the float64 operator, stores r
 * rounded to float32 and ||r|| (and ||b|| when first != 0) in the state;
 * esr_ir_update: x += (double) d, then d = 0; esr_ir_read: {rnorm, bnorm}. */
int esr_ir_workspace(int64_t n, int64_t *state_bytes, int64_t *ws_bytes);
int esr_ir_residual(int64_t n, const int32_t *rowptr, const int32_t *col,
                    const double *Ae, const double *Av, const double *b,
                    const double *x, float *r32, void *state, void *ws,
                    int first, void *stream);
int esr_ir_update(int64_t n, double *x, float *d, void *stream);
int esr_ir_read(const void *state, double *scalars, void *stream);

/* ---- batched Jacobi-PCG (new; the reference solves one system per CG) -----
 * cg.py:12-137 on K independent float64 Linsys instances concatenated into
 * one CSR (global rows and columns). Tiles of 256*R rows (R in 1, 2, 4, 8)
 * never straddle an instance: tinst/trow0/trow1 per tile, ifirst/intiles per
 * instance. state: K records of esr_bcg_state_bytes() bytes whose rtol, atol
 * and maxiter the caller sets before esr_bcg_init; `active` (int32) counts
 * instances not yet converged (the caller sets it to K) and each converged
 * instance's workgroups skip their tiles. part: 2 * ntiles doubles.       */
int esr_bcg_state_bytes(int64_t *bytes);
int esr_bcg_init(int R, int64_t ntiles, const int32_t *tinst,
                 const int32_t *trow0, const int32_t *trow1,
                 const int32_t *ifirst, const int32_t *intiles,
                 const int32_t *rowptr, const int32_t *col, const double *Ae,
                 const double *Av, const double *diag, const double *b,
                 double *x, double *r, double *z, double *p, void *state,
                 double *part, void *stream);
int esr_bcg_steps(int R, int64_t ntiles, const int32_t *tinst,
                  const int32_t *trow0, const int32_t *trow1,
                  const int32_t *ifirst, const int32_t *intiles,
                  const int32_t *rowptr, const int32_t *col, const double *Ae,
                  const double *Av, const double *diag, double *x, double *r,
                  double *z, double *p, double *Ap, void *state, double *part,
                  int32_t *active, int64_t count, void *stream);

/* ---- K <= 64 right-hand sides of one operator (cg.py:12-137 per
 * right-hand side, new form of the batched solver): vectors (n, K) row-major,
 * per-right-hand-side state (the batched solver's layout, rtol/atol/maxiter
 * set by the caller), the operator read once per iteration.                 */
int esr_mrhs_sizes(int64_t n, int64_t K, int64_t *state_bytes, int64_t *ws_bytes);
int esr_mrhs_init(int64_t n, int64_t K, const int32_t *rowptr, const int32_t *col,
                  const double *Ae, const double *Av, const double *diag,
                  const double *B, const double *X, double *R, double *Z, double *P,
                  void *state, double *ws, int32_t *active, void *stream);
int esr_mrhs_steps(int64_t n, int64_t K, const int32_t *rowptr, const int32_t *col,
                   const double *Ae, const double *Av, const double *diag, double *X,
                   double *R, double *Z, double *P, double *AP, void *state, double *ws,
                   int32_t *active, int64_t count, void *stream);
/* The same steps on a padded row-major ELL copy of the operator's
 * off-diagonal part, for operators whose rows hold at most ESR_MRHS_ELL_W
 * entries (the caller checks): len (n, uint8), ecol (n * W, int32; padding
 * = the row's own index), eval (n * W, float64; padding 0). esr_mrhs_ell
 * builds it from the CSR. Bit-identical to esr_mrhs_steps (same folds, same
 * order); the row's entries are loaded without waiting on row pointers, and
 * the next row's while this row's neighbours are gathered.               */
#define ESR_MRHS_ELL_W 8
int esr_mrhs_ell(int64_t n, const int32_t *rowptr, const int32_t *col,
                 const double *Ae, uint8_t *len, int32_t *ecol, double *eval,
                 void *stream);
int esr_mrhs_steps_ell(int64_t n, int64_t K, const uint8_t *len,
                       const int32_t *ecol, const double *eval, const double *Av,
                       const double *diag, double *X, double *R, double *Z,
                       double *P, double *AP, void *state, double *ws,
                       int32_t *active, int64_t count, void *stream);

/* ---- learned preconditioner (new; the reference's preconditioner slot,
 * easier/numeric/preconditioner/__init__.py:4, is empty) --------------------
 * z = M(r), an edge message-passing network with L layers of width 32 (see
 * easier_amd/numeric/preconditioner/learned.py for the flat parameter
 * layout). esr_gnn_sizes: float32 parameter count and bfloat16 packed-matrix
 * count; esr_gnn_pack: params -> packed (bf16 transposed Wp, Wq, Wu per
 * layer, the matrix-core operands); esr_gnn_forward: h holds 2 (or L+1 when
 * save) n x 32 bf16 buffers, Q 2 (or L), a L float32 n x 32 buffers when
 * save (else NULL), rnorm = device ||r||; esr_gnn_backward: gradient of
 * sum(dz * z) w.r.t. the parameters after a save=1 forward (rev = reverse
 * edge of every edge, work = float32 scratch, size in the source);
 * esr_gnn_adam: one Adam step (bias-corrected, optional gradient clip).     */
int esr_gnn_sizes(int L, int64_t *n_params, int64_t *n_packed);
int esr_gnn_pack(int L, const float *params, void *packed, void *stream);
int esr_gnn_forward(int L, int64_t n, const int32_t *rowptr,
                    const int32_t *col, const float *efeat, const float *vfeat,
                    const double *r, const double *d, const double *rnorm,
                    const float *params, const void *packed, void *h, void *Q,
                    float *a, int save, double *z, void *stream);
/* esr_gnn_forward with the edges slot-major per group of 16 vertices (one
 * wavefront's): edge k of vertex 16q+j at slot soff[q] + 16k + j (soff:
 * ceil(n/16)+1 int64), ecnt the degrees (uint8), ecol / efeat (3 floats) per
 * slot, padded slots with a valid column and zero features; the same
 * results as esr_gnn_forward on the CSR they are built from.                */
int esr_gnn_forward_slab(int L, int64_t n, const int64_t *soff,
                         const uint8_t *ecnt, const int32_t *ecol,
                         const float *efeat, const float *vfeat,
                         const double *r, const double *d, const double *rnorm,
                         const float *params, const void *packed, void *h,
                         void *Q, float *a, int save, double *z, void *stream);
int esr_gnn_backward(int L, int64_t n, int64_t nnz, const int32_t *rowptr,
                     const int32_t *col, const int32_t *rev,
                     const float *efeat, const float *vfeat, const double *r,
                     const double *d, const double *rnorm,
                     const float *params, const void *packed, const void *h,
                     const void *Q, const float *a, const double *dz,
                     float *grad, float *work, int64_t work_floats,
                     void *stream);
int esr_gnn_adam(int64_t np_, float *params, const float *grad, float *m,
                 float *v, double lr, double b1, double b2, double eps,
                 int64_t step, double clip, void *stream);

/* ---- flexible CG (Polak-Ribiere) steps for a varying preconditioner ------
 * op: 0 q = A p, p.q, alpha = rz / p.q; 1 r_old = r, r -= alpha q,
 * x += alpha p, ||r||^2; 2 z.r, z.r_old, beta (a0 != 0: restart, beta = 0),
 * rz = z.r -- a z with z.r <= 0 or not finite is replaced by r / d for the
 * step; 3 p = z + beta p (a0 != 0: p = z); 4 r = b - A x, ||r||^2, ||b||^2;
 * 5 z = r / d.
 * esr_fcg_set arms the step operations (0, 1, 2, 3, 5): they do nothing once
 * a step has brought ||r||^2 to tol2 or maxit steps were taken (maxit < 0:
 * no limit), so chunks of steps can be captured in a graph and replayed.
 * esr_fcg_read: {rz, pq, alpha, rr, zr, zro, beta, rr_true, bb} and
 * (ints nullable) {steps since esr_fcg_set, done, non-descent z.r <= 0}.   */
int esr_fcg_workspace(int64_t n, int64_t *state_bytes, int64_t *ws_bytes);
int esr_fcg_op(int op, int64_t n, const int32_t *rowptr, const int32_t *col,
               const double *Ae, const double *Av, const double *d,
               const double *b, double *x, double *r, double *ro, double *p,
               double *q, double *z, int a0, void *state, void *ws,
               void *stream);
/* esr_fcg_op with the operator of ops 0 (product) and 4 (residual) in its
 * SELL-64 form; the same results.                                           */
int esr_fcg_op_sell(int op, int64_t n, const int64_t *soff, const uint8_t *cnt,
                    const void *colv, int32_t idx16, const double *val,
                    const double *Av, const double *d, const double *b,
                    double *x, double *r, double *ro, double *p, double *q,
                    double *z, int a0, void *state, void *ws, void *stream);
int esr_fcg_set(void *state, double tol2, int64_t maxit, void *stream);
int esr_fcg_read(const void *state, double *scalars, int64_t *ints, void *stream);

/* ---- geometric multigrid on the 2-triangle mesh (new; a preconditioner for
 * the reference's callable-M slot, numeric/solver/cg.py:86 and
 * easier/numeric/preconditioner/__init__.py:4) ------------------------------
 * Grids of n cells per side, (n+1)^2 vertices; coarse levels lexicographic
 * (vertex (I, J) at J (n+1) + I), the finest level in any order given by
 * fpos (index -> position) and fidx (position -> index), both NULL for
 * lexicographic. P = P1 interpolation between nested meshes without the
 * coarse boundary columns; restriction = P^T.
 * esr_mg_pattern: vertex graph of the n-grid (rowptr (n+1)^2+1, col).
 * esr_mg_galerkin: coarse Ae/Av = P^T A P on that pattern (boundary rows
 *   identity); *bad counts fine couplings outside the coarse stencil (0 for
 *   an operator on the 2-triangle mesh's graph).
 * esr_mg_features: the learned smoother's inputs on a lexicographic level.
 * esr_mg_parents: parent map of a fine level (2 int32 per vertex: coarse
 *   parents, -1 = boundary/none, second = -2 for a single parent), read by
 *   the kernels that apply P instead of dividing grid positions.
 * esr_mg_restrict: rc = P^T t. esr_mg_prolong: z += P zc.
 * esr_mg_jacobi_pre: z = wd r, t = r - A z (wd = w / d per row).
 * esr_mg_jacobi_post: y = z + P zc (y = z when zc is NULL),
 *   out = y + wd (r - A y).
 * esr_mg_residual: t = r - A y, y = z (zc NULL) or y = z + P zc (stored).
 * esr_mg_dense: z = Ainv r (coarsest grid, Ainv row-major float64).
 * dtype ESR_F64 or ESR_F32 (vectors and the operator's Ae/Av).           */
int esr_mg_pattern(int64_t n, int32_t *rowptr, int32_t *col, void *stream);
int esr_mg_galerkin(int64_t nf, const int32_t *fidx, const int32_t *fpos,
                    const int32_t *f_rowptr, const int32_t *f_col,
                    const double *f_Ae, const double *f_Av,
                    const int32_t *c_rowptr, double *c_Ae, double *c_Av,
                    int32_t *bad, void *stream);
int esr_mg_features(int64_t n, const int32_t *rowptr, const int32_t *col,
                    const double *Ae, const double *Av, const double *dsum,
                    const float *logk, float *vfeat, float *efeat, void *stream);
int esr_mg_parents(int64_t nf, const int32_t *fpos, int32_t *par, void *stream);
int esr_mg_restrict(int dtype, int64_t nf, const int32_t *fidx, const void *t,
                    void *rc, void *stream);
int esr_mg_prolong(int dtype, int64_t nf, const int32_t *par, const void *zc,
                   void *z, void *stream);
int esr_mg_jacobi_pre(int dtype, int64_t n, const int32_t *rowptr,
                      const int32_t *col, const void *Ae, const void *Av,
                      const void *wd, const void *r, void *z, void *t, void *stream);
int esr_mg_jacobi_post(int dtype, int64_t nf, const int32_t *par,
                       const int32_t *rowptr, const int32_t *col,
                       const void *Ae, const void *Av, const void *wd,
                       const void *r, const void *z, const void *zc, void *out,
                       void *stream);
int esr_mg_residual(int dtype, int64_t n, int64_t nf, const int32_t *par,
                    const int32_t *rowptr, const int32_t *col, const void *Ae,
                    const void *Av, const void *r, const void *z, const void *zc,
                    void *y, void *t, void *stream);
/* The three sweeps above with the level operator in its SELL-64 form
 * (esr_sell_plan / esr_sell_fill; values in the sweep's dtype): the same
 * results. No reference counterpart (multigrid is this package's own). */
int esr_mg_jacobi_pre_sell(int dtype, int64_t n, const int64_t *soff,
                           const uint8_t *cnt, const void *colv, int32_t idx16,
                           const void *val, const void *Av, const void *wd,
                           const void *r, void *z, void *t, void *stream);
int esr_mg_jacobi_post_sell(int dtype, int64_t nf, const int32_t *par,
                            const int64_t *soff, const uint8_t *cnt,
                            const void *colv, int32_t idx16, const void *val,
                            const void *Av, const void *wd, const void *r,
                            const void *z, const void *zc, void *out,
                            void *stream);
int esr_mg_residual_sell(int dtype, int64_t n, int64_t nf, const int32_t *par,
                         const int64_t *soff, const uint8_t *cnt,
                         const void *colv, int32_t idx16, const void *val,
                         const void *Av, const void *r, const void *z,
                         const void *zc, void *y, void *t, void *stream);
int esr_mg_dense(int dtype, int64_t n, const double *Ainv, const void *r, void *z,
                 void *stream);

/* ---- algebraic multigrid by smoothed aggregation (new; the same
 * preconditioner slot, numeric/solver/cg.py:86, for an operator on any mesh)
 * esr_amg_coarsen (host): one level's set-up from the Linsys form of an SPD
 *   operator (host arrays; Av > 0): strength |a_ij| >= theta sqrt(a_ii a_jj),
 *   aggregates in three passes, P = (I - 4/(3 rho) D^-1 A) P_tent with rho
 *   the Gershgorin bound of D^-1 A (returned), R = P^T and the Galerkin
 *   operator P^T A P in the Linsys form. *plan owns the result;
 *   sizes = {nc, nnz(P), off-diagonal nnz of P^T A P}; nc = 0 when no vertex
 *   has a strong neighbour. nthreads host threads (deterministic results).
 * esr_amg_fetch (host): copies the plan into caller arrays (any may be
 *   NULL): agg (n, -1 = not aggregated), P (n+1, nnz(P), nnz(P)),
 *   R (nc+1, nnz(P), nnz(P)), coarse Linsys (nc+1, nnz, nnz, nc).
 * esr_amg_destroy (host): frees the plan.
 * esr_amg_restrict: rc = R t. esr_amg_prolong: z += P zc (CSR transfers,
 *   values in the dtype).
 * esr_amg_sweep: damped-Jacobi sweeps of a CSR level (Linsys form): mode 0
 *   z = wd r, out = r - A z; mode 1 out = z + wd (r - A z); mode 2
 *   out = r - A z (wd unused, may be NULL).
 *   These three spread each row over `lanes` lanes (1, 2, 4, ..., 64):
 *   lane l folds entries l, l + lanes, ... and the partial sums are combined
 *   by an xor butterfly (deterministic; lanes = 1 is the plain row fold).
 * esr_amg_jacobi_post_sell: mode 1 on the level's SELL-64 form (one row per
 *   thread, the fold of esr_mg_jacobi_post).                                */
int esr_amg_coarsen(int64_t n, const int32_t *rowptr, const int32_t *col,
                    const double *Ae, const double *Av, double theta, int32_t nthreads,
                    void **plan, int64_t *sizes, double *rho);
int esr_amg_fetch(void *plan, int32_t *agg, int32_t *p_rowptr, int32_t *p_col,
                  double *p_val, int32_t *r_rowptr, int32_t *r_col, double *r_val,
                  int32_t *c_rowptr, int32_t *c_col, double *c_Ae, double *c_Av);
int esr_amg_destroy(void *plan);
int esr_amg_restrict(int dtype, int64_t nc, const int32_t *r_rowptr,
                     const int32_t *r_col, const void *r_val, const void *t,
                     void *rc, int32_t lanes, void *stream);
int esr_amg_prolong(int dtype, int64_t nf, const int32_t *p_rowptr,
                    const int32_t *p_col, const void *p_val, const void *zc,
                    void *z, int32_t lanes, void *stream);
int esr_amg_sweep(int dtype, int32_t mode, int64_t n, const int32_t *rowptr,
                  const int32_t *col, const void *Ae, const void *Av,
                  const void *wd, const void *r, void *z, void *out,
                  int32_t lanes, void *stream);
int esr_amg_jacobi_post_sell(int dtype, int64_t n, const int64_t *soff,
                             const uint8_t *cnt, const void *colv, int32_t idx16,
                             const void *val, const void *Av, const void *wd,
                             const void *r, const void *z, void *out,
                             void *stream);

#ifdef __cplusplus
}
#endif
#endif /* ESR_HIP_H */
