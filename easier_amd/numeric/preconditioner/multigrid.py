"""Multilevel preconditioner z = V(r): a multigrid V(1,1) cycle with
damped-Jacobi or learned (message-passing network) smoothing, executed by
the kernels of csrc/k_mg.hip, on one of two hierarchies:

* geometric, for the n x n 2-triangle mesh (n = m 2^k, m <= 32): level 0 is
  the caller's operator in its own vertex order, mapped to grid positions
  through the vertex coordinates; level l+1 has half the cells per side,
  lexicographic order, and the Galerkin operator P^T A P (P = P1
  interpolation between the nested meshes), built on the device; the
  coarsening stops at n <= coarsest cells per side (or an odd n);
* algebraic (smoothed aggregation, csrc/host_amg.cpp), for an SPD operator
  on any mesh: aggregates of strongly coupled vertices, P = (I - 4/(3 rho)
  D^-1 A) P_tent, Galerkin coarse operators, set up on the host once per
  operator and moved to the device; the coarsening stops at `coarse_rows`
  rows.
The coarsest level is solved with its dense inverse.

No reference counterpart: the reference's preconditioner slot
(easier/numeric/preconditioner/__init__.py:4) is empty; its solvers take any
callable M (numeric/solver/cg.py:86, gmres.py:152). A local preconditioner
(Jacobi, or a network that sees 3 mesh hops) leaves the O(n^1/2) Krylov
iteration count of the Poisson problem unchanged; the coarse grids carry the
smooth error components across the mesh in one cycle.

Cycle, per level (r in, z out):
  Jacobi:  z0 = w r/d; t = r - A z0; zc = V(P^T t);
           y = z0 + P zc; z = y + w (r - A y)/d           (4 kernels; w/d
           precomputed per level)
  learned: z0 = S(r); t = r - A z0; zc = V(P^T t);
           then the Jacobi post-smoothing above (default), or
           y = z0 + P zc; t = r - A y; z = y + S(t)
with S the network (z = (r + s net(r/s, features))/d, s = rms(r)). The
learned cycle is not linear in r (S is positively homogeneous only), which
is why it runs inside flexible CG. The network smooths in one position and
damped Jacobi in the other: on the 2048 x 2048 variable-coefficient problem
that took 10 flexible-CG steps, the network in both positions 15. On an
algebraic hierarchy the network smooths the finest level only (the coarse
levels have no mesh geometry for its edge features).
"""
import ctypes
import os
import types
from typing import List, Optional

import numpy as np
import torch

from ... import _native as nat
from ...core import ops

_c_i32, _c_i64, _c_vp, _c_dbl = ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_double
nat.register_signatures({
    'esr_mg_pattern': (_c_i32, [_c_i64, _c_vp, _c_vp, _c_vp]),
    'esr_mg_galerkin': (_c_i32, [_c_i64] + [_c_vp] * 11),
    'esr_mg_features': (_c_i32, [_c_i64] + [_c_vp] * 9),
    'esr_mg_parents': (_c_i32, [_c_i64, _c_vp, _c_vp, _c_vp]),
    'esr_mg_restrict': (_c_i32, [_c_i32, _c_i64, _c_vp, _c_vp, _c_vp, _c_vp]),
    'esr_mg_prolong': (_c_i32, [_c_i32, _c_i64, _c_vp, _c_vp, _c_vp, _c_vp]),
    'esr_mg_jacobi_pre': (_c_i32, [_c_i32, _c_i64] + [_c_vp] * 9),
    'esr_mg_jacobi_post': (_c_i32, [_c_i32, _c_i64] + [_c_vp] * 11),
    'esr_mg_residual': (_c_i32, [_c_i32, _c_i64, _c_i64] + [_c_vp] * 11),
    'esr_mg_dense': (_c_i32, [_c_i32, _c_i64, _c_vp, _c_vp, _c_vp, _c_vp]),
    'esr_mg_jacobi_pre_sell': (_c_i32, [_c_i32, _c_i64] + [_c_vp] * 3 + [_c_i32] + [_c_vp] * 7),
    'esr_mg_jacobi_post_sell': (_c_i32, [_c_i32, _c_i64] + [_c_vp] * 4 + [_c_i32]
                                + [_c_vp] * 8),
    'esr_mg_residual_sell': (_c_i32, [_c_i32, _c_i64, _c_i64] + [_c_vp] * 4 + [_c_i32]
                             + [_c_vp] * 8),
    'esr_amg_coarsen': (_c_i32, [_c_i64] + [_c_vp] * 4 + [_c_dbl, _c_i32] + [_c_vp] * 3),
    'esr_amg_fetch': (_c_i32, [_c_vp] * 12),
    'esr_amg_destroy': (_c_i32, [_c_vp]),
    'esr_amg_restrict': (_c_i32, [_c_i32, _c_i64] + [_c_vp] * 5 + [_c_i32, _c_vp]),
    'esr_amg_prolong': (_c_i32, [_c_i32, _c_i64] + [_c_vp] * 5 + [_c_i32, _c_vp]),
    'esr_amg_sweep': (_c_i32, [_c_i32, _c_i32, _c_i64] + [_c_vp] * 8 + [_c_i32, _c_vp]),
    'esr_amg_jacobi_post_sell': (_c_i32, [_c_i32, _c_i64] + [_c_vp] * 3 + [_c_i32]
                                 + [_c_vp] * 7),
})


def amg_coarsen(rowptr, col, Ae, Av, theta=0.08, nthreads=None):
    """One smoothed-aggregation level of an SPD operator in the Linsys form
    (host numpy arrays), by the library's host code (csrc/host_amg.cpp).
    Returns dict(nc, rho, agg, P=(rowptr, col, val) by fine row,
    R=(rowptr, col, val) = P^T by coarse row, A=(rowptr, col, Ae, Av) the
    Galerkin operator P^T A P); nc = 0 when nothing aggregates."""
    n = int(np.shape(Av)[0])
    rp = np.ascontiguousarray(rowptr, np.int32)
    cl = np.ascontiguousarray(col, np.int32)
    ae = np.ascontiguousarray(Ae, np.float64)
    av = np.ascontiguousarray(Av, np.float64)
    if rp.shape[0] != n + 1 or cl.shape[0] != rp[-1] or ae.shape[0] != cl.shape[0]:
        raise ValueError('amg_coarsen: inconsistent Linsys arrays')
    if nthreads is None:
        nthreads = max(1, min(16, os.cpu_count() or 1))
    plan = ctypes.c_void_p()
    sizes = (ctypes.c_int64 * 3)()
    rho = ctypes.c_double()
    nat.call('esr_amg_coarsen', n, rp.ctypes.data, cl.ctypes.data, ae.ctypes.data,
             av.ctypes.data, float(theta), int(nthreads), ctypes.byref(plan),
             ctypes.byref(sizes), ctypes.byref(rho))
    try:
        nc, nnzp, nnzc = (int(v) for v in sizes)
        out = dict(nc=nc, rho=rho.value, agg=np.empty(n, np.int32))
        a = dict(prow=np.empty(n + 1 if nc else 0, np.int32), pcol=np.empty(nnzp, np.int32),
                 pval=np.empty(nnzp), rrow=np.empty(nc + 1 if nc else 0, np.int32),
                 rcol=np.empty(nnzp, np.int32), rval=np.empty(nnzp),
                 crow=np.empty(nc + 1 if nc else 0, np.int32), ccol=np.empty(nnzc, np.int32),
                 cAe=np.empty(nnzc), cAv=np.empty(nc))
        nat.call('esr_amg_fetch', plan, out['agg'].ctypes.data,
                 *[a[k].ctypes.data for k in ('prow', 'pcol', 'pval', 'rrow', 'rcol', 'rval',
                                              'crow', 'ccol', 'cAe', 'cAv')])
    finally:
        nat.call('esr_amg_destroy', plan)
    out['P'] = (a['prow'], a['pcol'], a['pval'])
    out['R'] = (a['rrow'], a['rcol'], a['rval'])
    out['A'] = (a['crow'], a['ccol'], a['cAe'], a['cAv'])
    return out


def row_lanes(nnz, rows):
    """Lanes per row of the algebraic-level kernels: about two entries per
    lane on large levels (the power of two nearest to half the mean row
    length: memory-level parallelism without idle lanes), one per lane on
    small ones (the row's load chain is the whole cost). At 2048^2 the
    flexible-CG solve took 12.70 ms with 8 lanes for the restriction into
    level 1 (15 entries per row), 13.11 ms with 4, 15.41 ms with 2; the
    level-1 sweeps (8 entries per row) were fastest with 4 lanes."""
    avg = max(float(nnz) / max(int(rows), 1), 1.0)
    if rows >= 65536:
        lanes = 1 << int(np.round(np.log2(max(avg / 2.0, 1.0))))
    else:
        lanes = 1 << int(np.ceil(np.log2(avg)))
    return int(min(max(lanes, 1), 64))


def grid_positions(points, n):
    """Lexicographic grid position J (n+1) + I of each vertex (I = rint(x n),
    J = rint(y n)): exact for the 2-triangle mesh on the unit square, with or
    without interior jitter below half a cell. Raises ValueError unless the
    vertices are the (n+1)^2 grid points, each once."""
    pts = np.asarray(points, np.float64)
    W = n + 1
    if pts.ndim != 2 or pts.shape[0] != W * W:
        raise ValueError(f'{pts.shape[0]} vertices are not the vertices of a {n} x {n} grid')
    I = np.rint(pts[:, 0] * n).astype(np.int64)
    J = np.rint(pts[:, 1] * n).astype(np.int64)
    if I.min() < 0 or J.min() < 0 or I.max() > n or J.max() > n:
        raise ValueError('vertex coordinates outside the unit square')
    pos = J * W + I
    if np.bincount(pos, minlength=W * W).max() != 1:
        raise ValueError('two vertices map to one grid position')
    return pos


def levels_for(n, coarsest=16):
    """Grid sizes of the hierarchy: n, n/2, ... down to n <= coarsest or odd."""
    out = [int(n)]
    while out[-1] % 2 == 0 and out[-1] > coarsest:
        out.append(out[-1] // 2)
    return out


class _Graph:
    """The network's static inputs on a coarse level (inference only)."""

    def __init__(self, n, rowptr, col, efeat, vfeat, d):
        self.n, self.rowptr, self.col, self.efeat, self.vfeat, self.d = \
            n, rowptr, col, efeat, vfeat, d
        self.slab = None


class Level:
    """One level: its operator (Linsys form, device), and the transfers into
    the next coarser level -- geometric (grid size n_grid, parent map par)
    or algebraic (CSR P by fine row and R = P^T by coarse row)."""

    def __init__(self, n_grid, rowptr, col, Ae, Av, fpos=None, fidx=None, nv=None):
        self.n_grid = n_grid
        self.nv = (n_grid + 1) ** 2 if nv is None else int(nv)
        self.rowptr, self.col, self.Ae, self.Av = rowptr, col, Ae, Av
        self.fpos, self.fidx = fpos, fidx
        self.Ainv = None
        self.par = None          # parent map into the next coarser level
        self.P = self.R = None   # algebraic transfers (rowptr, col, val)
        self.rho = None          # Gershgorin bound of D^-1 A (algebraic levels)
        self.graph = None        # network inputs (learned smoothing)
        self.bufs = None


class MultigridPreconditioner:
    """V(1,1)-cycle preconditioner. cycle_dtype float32 runs the whole cycle
    (vectors, level operators, sweeps and transfers) in float32 on float32
    copies of the Galerkin operators, built in float64: the mixed-precision
    form in which flexible CG keeps dot products, the iteration and the
    convergence test in float64. smoother: None (damped Jacobi, weight
    omega) or a LearnedPreconditioner (its network smooths every level with at
    least `learned_min_grid` cells per side, before the coarse-grid correction
    if learned_pre and after it if learned_post, damped Jacobi otherwise;
    coarser levels use Jacobi).

    coarsening: 'geometric' (the n x n 2-triangle mesh, n = m 2^k, m <= 32),
    'aggregation' (smoothed aggregation of the operator itself: any mesh,
    any SPD Linsys; strength threshold theta, halved per level; coarsening
    stops at `coarse_rows` rows, solved by the dense inverse; Jacobi weight
    omega * 2 / rho per level, rho the Gershgorin bound of D^-1 A; a learned
    smoother smooths the finest level only), or 'auto' (geometric where it
    applies, else aggregation).
    setup() builds the hierarchy for one operator; apply(r, z) runs one cycle
    and allocates nothing, so it can be captured in a HIP graph."""

    _bindings = 0
    fcg_chunk = 4        # flexible-CG steps per captured graph

    def __init__(self, smoother=None, omega=0.8, coarsest=16, learned_min_grid=64,
                 learned_pre=True, learned_post=False, cycle_dtype=torch.float64,
                 coarsening='auto', theta=0.08, coarse_rows=1024, max_levels=25):
        if cycle_dtype not in (torch.float64, torch.float32):
            raise ValueError('cycle_dtype must be float64 or float32')
        if smoother is not None and cycle_dtype != torch.float64:
            raise ValueError('the learned smoother runs in float64 cycles only')
        if coarsening not in ('auto', 'geometric', 'aggregation'):
            raise ValueError("coarsening must be 'auto', 'geometric' or 'aggregation'")
        self.smoother = smoother
        self.cycle_dtype = cycle_dtype
        self.dt = nat.dtype_code(cycle_dtype)
        self.learned_pre, self.learned_post = bool(learned_pre), bool(learned_post)
        self.omega = float(omega)
        self.coarsest = int(coarsest)
        self.learned_min_grid = int(learned_min_grid)
        self.coarsening = coarsening
        self.theta = float(theta)
        self.coarse_rows = int(coarse_rows)
        self.max_levels = int(max_levels)
        self.kind = None         # the hierarchy set up: 'geometric' or 'aggregation'
        self.setup_seconds = {}
        self.levels: List[Level] = []
        self.binding = 0

    # -- set-up ---------------------------------------------------------------
    def geometric_unfit(self, nv, n_grid):
        """Why the geometric hierarchy does not apply to an operator of nv rows
        on an n_grid mesh (None when it does)."""
        if n_grid is None:
            return 'no grid size given'
        if nv != (n_grid + 1) ** 2:
            return f'operator of {nv} rows is not on a {n_grid} x {n_grid} grid'
        sizes = levels_for(n_grid, self.coarsest)
        if sizes[-1] > 32:
            return (f'grid {n_grid} coarsens to an odd grid of {sizes[-1]} cells per side: '
                    'the coarsest grid must have at most 32')
        return None

    def setup(self, rowptr, col, Ae, Av, n_grid=None, fpos=None, logk=None, graph0=None):
        """rowptr/col int32, Ae/Av float64 device tensors of the fine operator
        (Linsys form: off-diagonal CSR + diagonal). Geometric coarsening
        needs n_grid (the operator on the n_grid x n_grid 2-triangle mesh) and
        fpos (host int array, vertex -> grid position; None = lexicographic).
        With a learned smoother: logk (host, per vertex in the operator's
        order) and graph0, the network's inputs on the fine level (a
        MeshGraph)."""
        import time
        nv = Av.shape[0]
        kind = self.coarsening
        t0 = time.perf_counter()
        if kind != 'aggregation':
            why = self.geometric_unfit(nv, n_grid)
            if why is None:
                try:
                    self._setup_geometric(rowptr, col, Ae, Av, n_grid, fpos)
                except ValueError as e:
                    why = str(e)
            if why is not None:
                if kind == 'geometric':
                    raise ValueError(why)
                kind = 'aggregation'
            else:
                kind = 'geometric'
        if kind == 'aggregation':
            self._setup_aggregation(rowptr, col, Ae, Av)
        self.kind = kind
        self.setup_seconds['hierarchy'] = time.perf_counter() - t0
        dev = Av.device
        cd = self.cycle_dtype
        for k, L in enumerate(self.levels):
            # the operator the cycle reads (a float32 copy in a float32 cycle)
            L.cAe = L.Ae if cd == torch.float64 else L.Ae.to(cd)
            L.cAv = L.Av if cd == torch.float64 else L.Av.to(cd)
            if L.P is not None:
                L.cP = (L.P[0], L.P[1], L.P[2].to(cd))
                L.cR = (L.R[0], L.R[1], L.R[2].to(cd))
                C = self.levels[k + 1]
                L.lanes_P = row_lanes(L.P[1].shape[0], L.nv)
                L.lanes_R = row_lanes(L.R[1].shape[0], C.nv)
            # lanes per row of a CSR sweep of an algebraic level: the finest
            # keeps the one-row-per-thread fold of its SELL form
            L.lanes_A = 1 if k == 0 else row_lanes(L.col.shape[0], L.nv)
            L.sell = None
            from ..solver.fused import FusedCG, make_sell
            # coarse algebraic levels: rows spread over lanes instead
            if L.Ainv is None and FusedCG.use_sell and (self.kind == 'geometric' or k == 0):
                L.sell = make_sell(types.SimpleNamespace(rowptr=L.rowptr, col=L.col, Ae=L.cAe,
                                                         Av=L.cAv, n=L.nv))
            w = self.omega if L.rho is None else self.omega * 2.0 / L.rho
            L.wd = ops.ewise('div', w, L.Av)                  # w / d, float64
            if cd != torch.float64:
                L.wd = L.wd.to(cd)
            L.z = torch.zeros(L.nv, dtype=cd, device=dev)
            L.t = torch.zeros(L.nv, dtype=cd, device=dev)
            if k > 0:
                L.r = torch.zeros(L.nv, dtype=cd, device=dev)
                L.zo = torch.zeros(L.nv, dtype=cd, device=dev)
        if cd != torch.float64:
            self.r_in = torch.zeros(nv, dtype=cd, device=dev)
            self.z_out = torch.zeros(nv, dtype=cd, device=dev)
        if self.smoother is not None:
            self._setup_learned(logk, graph0, fpos)
        MultigridPreconditioner._bindings += 1
        self.binding = MultigridPreconditioner._bindings
        self.setup_seconds['total'] = time.perf_counter() - t0
        return self

    def _setup_geometric(self, rowptr, col, Ae, Av, n_grid, fpos):
        dev = Av.device
        nv = Av.shape[0]
        sizes = levels_for(n_grid, self.coarsest)
        i32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(dev)
        fp = fi = None
        if fpos is not None:
            fpos = np.asarray(fpos, np.int64)
            if not np.array_equal(fpos, np.arange(nv)):
                fidx = np.empty_like(fpos)
                fidx[fpos] = np.arange(nv)
                fp, fi = i32(fpos), i32(fidx)
        levels = [Level(n_grid, rowptr, col, Ae, Av, fp, fi)]
        bad = torch.zeros(1, dtype=torch.int32, device=dev)
        sp = nat.stream_ptr()
        for nc in sizes[1:]:
            f = levels[-1]
            W = nc + 1
            c_rowptr = torch.empty(W * W + 1, dtype=torch.int32, device=dev)
            c_col = torch.empty((W - 1) * (6 * W - 2), dtype=torch.int32, device=dev)
            nat.call('esr_mg_pattern', nc, c_rowptr.data_ptr(), c_col.data_ptr(), sp)
            c_Ae = torch.empty(c_col.shape[0], dtype=torch.float64, device=dev)
            c_Av = torch.empty(W * W, dtype=torch.float64, device=dev)
            P = lambda t: None if t is None else t.data_ptr()
            nat.call('esr_mg_galerkin', f.n_grid, P(f.fidx), P(f.fpos), f.rowptr.data_ptr(),
                     f.col.data_ptr(), f.Ae.data_ptr(), f.Av.data_ptr(), c_rowptr.data_ptr(),
                     c_Ae.data_ptr(), c_Av.data_ptr(), bad.data_ptr(), sp)
            levels.append(Level(nc, c_rowptr, c_col, c_Ae, c_Av))
        for L in levels[:-1]:
            L.par = torch.empty(2 * L.nv, dtype=torch.int32, device=dev)
            nat.call('esr_mg_parents', L.n_grid, None if L.fpos is None else L.fpos.data_ptr(),
                     L.par.data_ptr(), sp)
        if int(bad.item()):
            raise ValueError('the operator couples vertices that are not neighbours on the '
                             '2-triangle mesh: geometric coarsening does not apply')
        last = levels[-1]
        last.Ainv = self._dense_inverse(last.rowptr.cpu().numpy(), last.col.cpu().numpy(),
                                        last.Ae.cpu().numpy(), last.Av.cpu().numpy(), dev)
        self.levels = levels

    def _setup_aggregation(self, rowptr, col, Ae, Av):
        dev = Av.device
        i32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(dev)
        f64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)
        levels = [Level(None, rowptr, col, Ae, Av, nv=Av.shape[0])]
        host = (rowptr.cpu().numpy(), col.cpu().numpy(), Ae.cpu().numpy(), Av.cpu().numpy())
        while levels[-1].nv > self.coarse_rows and len(levels) < self.max_levels:
            L = levels[-1]
            lv = amg_coarsen(*host, theta=self.theta * 0.5 ** (len(levels) - 1))
            L.rho = lv['rho']
            if lv['nc'] == 0 or lv['nc'] > 0.9 * L.nv:
                break
            L.P = tuple(f(a) for f, a in zip((i32, i32, f64), lv['P']))
            L.R = tuple(f(a) for f, a in zip((i32, i32, f64), lv['R']))
            crow, ccol, cAe, cAv = lv['A']
            levels.append(Level(None, i32(crow), i32(ccol), f64(cAe), f64(cAv), nv=lv['nc']))
            host = lv['A']
        last = levels[-1]
        if last.nv <= 8 * self.coarse_rows:
            last.Ainv = self._dense_inverse(*host, dev)
        # else: too few strong couplings to coarsen further and too large to
        # invert -- the last level is smoothed only (two damped-Jacobi sweeps)
        self.levels = levels

    @staticmethod
    def _dense_inverse(rp, cl, ae, av, dev):
        nv = av.shape[0]
        dense = np.zeros((nv, nv))
        rows = np.repeat(np.arange(nv), np.diff(rp))
        np.add.at(dense, (rows, cl), ae)
        dense[np.arange(nv), np.arange(nv)] += av
        return torch.from_numpy(np.linalg.inv(dense)).to(dev)

    def _setup_learned(self, logk, graph0, fpos):
        from .learned import EdgeSlab, GnnBuffers
        if logk is None or graph0 is None:
            raise ValueError('a learned smoother needs logk and the fine-level MeshGraph')
        dev = self.levels[0].Av.device
        if self.kind == 'aggregation':
            # coarse algebraic levels have no mesh geometry: the finest level only
            L = self.levels[0]
            if L.Ainv is None:
                L.graph = graph0
                L.bufs = GnnBuffers(L.nv, dev)
                L.s = torch.zeros(L.nv, dtype=torch.float64, device=dev)
            return
        n0 = self.levels[0].n_grid
        W0 = n0 + 1
        lk = np.asarray(logk, np.float64).reshape(-1)
        lex = np.empty(W0 * W0)
        lex[np.arange(W0 * W0) if fpos is None else np.asarray(fpos)] = lk
        lex = lex.reshape(W0, W0)
        dsum = torch.empty(1, dtype=torch.float64, device=dev)
        for k, L in enumerate(self.levels):
            if L.Ainv is not None or L.n_grid < self.learned_min_grid:
                continue
            if k == 0:
                L.graph = graph0
            else:
                step = 2 ** k
                lkl = torch.from_numpy(np.ascontiguousarray(lex[::step, ::step],
                                                            np.float32).reshape(-1)).to(dev)
                vfeat = torch.empty((L.nv, 2), dtype=torch.float32, device=dev)
                efeat = torch.empty((L.col.shape[0], 3), dtype=torch.float32, device=dev)
                ops.aggregate('sum', L.Av.reshape(-1, 1), out=dsum)
                nat.call('esr_mg_features', L.n_grid, L.rowptr.data_ptr(), L.col.data_ptr(),
                         L.Ae.data_ptr(), L.Av.data_ptr(), dsum.data_ptr(), lkl.data_ptr(),
                         vfeat.data_ptr(), efeat.data_ptr(), nat.stream_ptr())
                L.graph = _Graph(L.nv, L.rowptr, L.col, efeat, vfeat, L.Av)
                L.graph.slab = EdgeSlab.of(L.graph)
            L.bufs = GnnBuffers(L.nv, dev)
            L.s = torch.zeros(L.nv, dtype=torch.float64, device=dev)

    # -- cycle ----------------------------------------------------------------
    def apply(self, r: torch.Tensor, z: torch.Tensor):
        if not self.levels:
            raise RuntimeError('MultigridPreconditioner.setup() first')
        if self.cycle_dtype == torch.float64:
            self._cycle(0, r, z)
        else:
            ops.ewise('copy', r, out=self.r_in)
            self._cycle(0, self.r_in, self.z_out)
            ops.ewise('copy', self.z_out, out=z)
        return z

    def __call__(self, r):
        return self.apply(r, torch.empty_like(r))

    # sweeps on a level's operator: its SELL form when it has one, else CSR
    def _sweep(self, L, mode, r, z, out, wd=True):
        """esr_amg_sweep on an algebraic level's CSR (rows over L.lanes_A lanes)."""
        nat.call('esr_amg_sweep', self.dt, mode, L.nv, L.rowptr.data_ptr(), L.col.data_ptr(),
                 L.cAe.data_ptr(), L.cAv.data_ptr(), L.wd.data_ptr() if wd else None,
                 r.data_ptr(), z.data_ptr(), out.data_ptr(), L.lanes_A, nat.stream_ptr())

    def _pre(self, L, r, z, t):
        sp, DT = nat.stream_ptr(), self.dt
        if L.sell is not None:
            nat.call('esr_mg_jacobi_pre_sell', DT, L.nv, *L.sell.arrays(), L.cAv.data_ptr(),
                     L.wd.data_ptr(), r.data_ptr(), z.data_ptr(), t.data_ptr(), sp)
        elif self.kind == 'aggregation':
            self._sweep(L, 0, r, z, t)
        else:
            nat.call('esr_mg_jacobi_pre', DT, L.nv, L.rowptr.data_ptr(), L.col.data_ptr(),
                     L.cAe.data_ptr(), L.cAv.data_ptr(), L.wd.data_ptr(), r.data_ptr(),
                     z.data_ptr(), t.data_ptr(), sp)

    def _residual(self, L, r, z, t):
        sp, DT = nat.stream_ptr(), self.dt
        if L.sell is not None:
            nat.call('esr_mg_residual_sell', DT, L.nv, L.n_grid or 0, None, *L.sell.arrays(),
                     L.cAv.data_ptr(), r.data_ptr(), z.data_ptr(), None, None, t.data_ptr(), sp)
        elif self.kind == 'aggregation':
            self._sweep(L, 2, r, z, t, wd=False)
        else:
            nat.call('esr_mg_residual', DT, L.nv, L.n_grid or 0, None, L.rowptr.data_ptr(),
                     L.col.data_ptr(), L.cAe.data_ptr(), L.cAv.data_ptr(), r.data_ptr(),
                     z.data_ptr(), None, None, t.data_ptr(), sp)

    def _restrict(self, L, C, t):
        """C.r = P^T t."""
        sp, DT = nat.stream_ptr(), self.dt
        if L.P is not None:
            nat.call('esr_amg_restrict', DT, C.nv, L.cR[0].data_ptr(), L.cR[1].data_ptr(),
                     L.cR[2].data_ptr(), t.data_ptr(), C.r.data_ptr(), L.lanes_R, sp)
        else:
            nat.call('esr_mg_restrict', DT, L.n_grid,
                     None if L.fidx is None else L.fidx.data_ptr(), t.data_ptr(),
                     C.r.data_ptr(), sp)

    def _prolong(self, L, C, z):
        """z += P C.zo."""
        sp, DT = nat.stream_ptr(), self.dt
        if L.P is not None:
            nat.call('esr_amg_prolong', DT, L.nv, L.cP[0].data_ptr(), L.cP[1].data_ptr(),
                     L.cP[2].data_ptr(), C.zo.data_ptr(), z.data_ptr(), L.lanes_P, sp)
        else:
            nat.call('esr_mg_prolong', DT, L.n_grid, L.par.data_ptr(), C.zo.data_ptr(),
                     z.data_ptr(), sp)

    def _cycle(self, k, r, out):
        L = self.levels[k]
        sp = nat.stream_ptr()
        DT = self.dt           # the cycle's dtype code
        if L.Ainv is not None:
            nat.call('esr_mg_dense', DT, L.nv, L.Ainv.data_ptr(), r.data_ptr(),
                     out.data_ptr(), sp)
            return
        if k + 1 == len(self.levels):
            # an algebraic level that could not be coarsened: smoothing only
            if L.graph is not None and self.learned_pre:
                self.smoother.forward_into(L.graph, L.bufs, r, L.z)
            else:
                self._pre(L, r, L.z, L.t)
            self._post_sweep(L, r, out)
            return
        C = self.levels[k + 1]
        if L.graph is None:
            self._pre(L, r, L.z, L.t)
            self._restrict(L, C, L.t)
            self._cycle(k + 1, C.r, C.zo)
            self._post(L, C, r, out)
            return
        S = self.smoother
        if self.learned_pre:
            S.forward_into(L.graph, L.bufs, r, L.z)
            self._residual(L, r, L.z, L.t)
        else:
            self._pre(L, r, L.z, L.t)
        self._restrict(L, C, L.t)
        self._cycle(k + 1, C.r, C.zo)
        if self.learned_post:
            self._prolong(L, C, L.z)
            self._residual(L, r, L.z, L.t)
            S.forward_into(L.graph, L.bufs, L.t, L.s)
            ops.ewise('add', L.z, L.s, out=out)
        else:
            self._post(L, C, r, out)

    def _post_sweep(self, L, r, out):
        """out = z + w (r - A z) / d on an algebraic level."""
        if L.sell is not None:
            nat.call('esr_amg_jacobi_post_sell', self.dt, L.nv, *L.sell.arrays(),
                     L.cAv.data_ptr(), L.wd.data_ptr(), r.data_ptr(), L.z.data_ptr(),
                     out.data_ptr(), nat.stream_ptr())
        else:
            self._sweep(L, 1, r, L.z, out)

    def _post(self, L, C, r, out):
        """z += P zc (in place), then out = z + w (r - A z) / d."""
        sp, DT = nat.stream_ptr(), self.dt
        self._prolong(L, C, L.z)
        if L.P is not None:
            self._post_sweep(L, r, out)
        elif L.sell is not None:
            nat.call('esr_mg_jacobi_post_sell', DT, L.n_grid, None, *L.sell.arrays(),
                     L.cAv.data_ptr(), L.wd.data_ptr(), r.data_ptr(), L.z.data_ptr(), None,
                     out.data_ptr(), sp)
        else:
            nat.call('esr_mg_jacobi_post', DT, L.n_grid, None, L.rowptr.data_ptr(),
                     L.col.data_ptr(), L.cAe.data_ptr(), L.cAv.data_ptr(), L.wd.data_ptr(),
                     r.data_ptr(), L.z.data_ptr(), None, out.data_ptr(), sp)
