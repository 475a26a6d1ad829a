"""Batched Jacobi-PCG: many independent Poisson systems (uneven meshes, or
many right-hand sides of one mesh) solved together.

No reference counterpart (the reference's CG solves one system). Each
instance follows cg.py:106-137 exactly (its own tol, iters, residual);
instances are concatenated into one CSR, processed by the same three kernels
per iteration (SpMV + p.Ap; x/r/z + r.r, r.z; test + p), with per-instance
device scalars and convergence masks: a converged instance's workgroups return
at once. The iteration loop is one captured HIP graph replayed until the
device counter of active instances reaches zero. The number of workgroups per
instance (about n_k / (256 R), R rows per thread) is picked by a small
autotuner that times a few iterations of each candidate R on the actual
batch.
"""
import ctypes
import time
from typing import Dict, List, Optional

import numpy as np
import torch

from ... import _native as nat
from ...core import ops
from .fused import _Graph

_c_i32, _c_i64, _c_vp = ctypes.c_int, ctypes.c_int64, ctypes.c_void_p
nat.register_signatures({
    'esr_bcg_state_bytes': (_c_i32, [ctypes.POINTER(_c_i64)]),
    'esr_bcg_init': (_c_i32, [_c_i32, _c_i64] + [_c_vp] * 17 + [_c_vp]),
    'esr_bcg_steps': (_c_i32, [_c_i32, _c_i64] + [_c_vp] * 18 + [_c_i64, _c_vp]),
    'esr_mrhs_sizes': (_c_i32, [_c_i64, _c_i64, ctypes.POINTER(_c_i64),
                                ctypes.POINTER(_c_i64)]),
    'esr_mrhs_init': (_c_i32, [_c_i64, _c_i64] + [_c_vp] * 14),
    'esr_mrhs_steps': (_c_i32, [_c_i64, _c_i64] + [_c_vp] * 13 + [_c_i64, _c_vp]),
    'esr_mrhs_ell': (_c_i32, [_c_i64] + [_c_vp] * 7),
    'esr_mrhs_steps_ell': (_c_i32, [_c_i64, _c_i64] + [_c_vp] * 13 + [_c_i64, _c_vp]),
})
ELL_W = 8            # ESR_MRHS_ELL_W

_STATE_DTYPE = np.dtype([
    ('rz_cur', 'f8'), ('rz_next', 'f8'), ('pAp', 'f8'), ('alpha', 'f8'),
    ('rr', 'f8'), ('rz_new', 'f8'), ('rnorm', 'f8'), ('bnorm', 'f8'),
    ('tol', 'f8'), ('rtol', 'f8'), ('atol', 'f8'), ('pad', 'f8'),
    ('k_cur', 'i8'), ('k_next', 'i8'), ('maxiter', 'i8'), ('done', 'i8'),
    ('iters', 'i8'), ('pad2', 'i8')])
_CANDIDATES = (1, 2, 4, 8)
_MAX_TILES = 256     # workgroups per instance (each sums its instance's partials)


def _same_operator(systems) -> bool:
    a = systems[0]
    for s in systems[1:]:
        for k in ('row', 'col', 'Ae', 'Av'):
            if s[k] is not a[k] and not (s[k].shape == a[k].shape and
                                         np.array_equal(s[k], a[k])):
                return False
    return True


class MultiRhsCG:
    """Jacobi-PCG for K <= 64 right-hand sides of one operator, the operator
    read once per iteration for all of them (csrc/k_mrhs.hip): vectors are
    (n, K), each right-hand side keeps its own scalars, tolerance and
    iteration count (cg.py:106-137 per right-hand side)."""

    MAX_K = 64
    use_ell = True       # the padded ELL product when every row fits ELL_W entries

    def __init__(self, system: Dict[str, np.ndarray], B: np.ndarray, jacobi=True,
                 device=None):
        dev = device or torch.device('cuda', torch.cuda.current_device())
        n = system['Av'].shape[0]
        B = np.ascontiguousarray(B, np.float64).reshape(n, -1)
        self.n, self.K = n, B.shape[1]
        if not 1 <= self.K <= self.MAX_K:
            raise ValueError(f'1 <= K <= {self.MAX_K} right-hand sides, got {self.K}')
        if np.any(np.diff(system['row']) < 0):
            raise ValueError('system rows must be sorted')
        rowptr = np.concatenate([[0], np.cumsum(np.bincount(system['row'], minlength=n))])
        t32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(dev)
        f64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)
        self.rowptr, self.col = t32(rowptr), t32(system['col'])
        self.Ae, self.Av = f64(system['Ae']), f64(system['Av'])
        self.diag = self.Av if jacobi else None
        self.ell = None
        deg = np.diff(rowptr)
        if self.use_ell and deg.size and int(deg.max()) <= ELL_W:
            self.ell = (torch.empty(n, dtype=torch.uint8, device=dev),
                        torch.empty(n * ELL_W, dtype=torch.int32, device=dev),
                        torch.empty(n * ELL_W, dtype=torch.float64, device=dev))
            nat.call('esr_mrhs_ell', n, self.rowptr.data_ptr(), self.col.data_ptr(),
                     self.Ae.data_ptr(), *[t.data_ptr() for t in self.ell], nat.stream_ptr())
        self.B = f64(B)
        mk = lambda: torch.zeros((n, self.K), dtype=torch.float64, device=dev)
        self.X, self.R, self.Z, self.P, self.AP = mk(), mk(), mk(), mk(), mk()
        sb, wb = ctypes.c_int64(), ctypes.c_int64()
        nat.call('esr_mrhs_sizes', n, self.K, ctypes.byref(sb), ctypes.byref(wb))
        assert sb.value == self.K * _STATE_DTYPE.itemsize
        self.state = torch.zeros(sb.value, dtype=torch.uint8, device=dev)
        self.ws = torch.empty(wb.value, dtype=torch.uint8, device=dev)
        self.active = torch.zeros(1, dtype=torch.int32, device=dev)
        self.dev = dev
        self.graphs = {}

    def _op(self):
        return [self.rowptr.data_ptr(), self.col.data_ptr(), self.Ae.data_ptr(),
                self.Av.data_ptr(), None if self.diag is None else self.diag.data_ptr()]

    def _steps(self, count, sp):
        vecs = [self.X.data_ptr(), self.R.data_ptr(), self.Z.data_ptr(), self.P.data_ptr(),
                self.AP.data_ptr(), self.state.data_ptr(), self.ws.data_ptr(),
                self.active.data_ptr(), count, sp]
        if self.ell is not None:
            nat.call('esr_mrhs_steps_ell', self.n, self.K, *[t.data_ptr() for t in self.ell],
                     self.Av.data_ptr(), None if self.diag is None else self.diag.data_ptr(),
                     *vecs)
        else:
            nat.call('esr_mrhs_steps', self.n, self.K, *self._op(), *vecs)

    def step_time(self, iters=5):
        """Milliseconds per iteration, from HIP events around `iters`
        iterations of a solve armed with tolerance 0 (every step runs)."""
        self._arm(0.0, None, None)
        self._steps(1, nat.stream_ptr())
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        self._steps(iters, nat.stream_ptr())
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / iters

    def _arm(self, rtol, atol, maxiter):
        st = np.zeros(self.K, _STATE_DTYPE)
        st['rtol'] = rtol
        st['atol'] = atol if atol else 0.0
        st['maxiter'] = -1 if maxiter is None else maxiter
        self.state.copy_(torch.from_numpy(st.view(np.uint8)).to(self.dev))
        ops.fill_(self.X, 0)
        nat.call('esr_mrhs_init', self.n, self.K, *self._op(), self.B.data_ptr(),
                 self.X.data_ptr(), self.R.data_ptr(), self.Z.data_ptr(), self.P.data_ptr(),
                 self.state.data_ptr(), self.ws.data_ptr(), self.active.data_ptr(),
                 nat.stream_ptr())

    def solve(self, rtol=1e-8, atol=None, maxiter=None, chunk=32):
        st = np.zeros(self.K, _STATE_DTYPE)
        st['rtol'] = rtol
        st['atol'] = atol if atol else 0.0
        st['maxiter'] = -1 if maxiter is None else maxiter
        self.state.copy_(torch.from_numpy(st.view(np.uint8)).to(self.dev))
        ops.fill_(self.X, 0)
        nat.call('esr_mrhs_init', self.n, self.K, *self._op(), self.B.data_ptr(),
                 self.X.data_ptr(), self.R.data_ptr(), self.Z.data_ptr(), self.P.data_ptr(),
                 self.state.data_ptr(), self.ws.data_ptr(), self.active.data_ptr(),
                 nat.stream_ptr())
        if chunk not in self.graphs:
            self.graphs[chunk] = _Graph(lambda sp: self._steps(chunk, sp))
        g = self.graphs[chunk]
        while True:
            g.replay()
            if int(self.active.item()) == 0:
                break
        st = self.state.cpu().numpy().view(_STATE_DTYPE)
        return [{'residual': float(st['rnorm'][k]), 'iters': int(st['iters'][k])}
                for k in range(self.K)]

    def solutions(self):
        X = self.X.cpu().numpy()
        return [np.ascontiguousarray(X[:, k]) for k in range(self.K)]


class BatchedCG:
    """systems: list of dicts with host arrays row, col (int64, CSR order:
    sorted by row), Ae, Av, b (float64), as workloads.poisson.p1_system
    returns. Solutions land in `x` (one float64 array per instance via
    solutions()). When every system has the same operator (right-hand sides
    of one mesh, at most 64), the shared-operator solver MultiRhsCG runs
    instead."""

    def __init__(self, systems: List[Dict[str, np.ndarray]], jacobi=True,
                 device=None, share_operator=True):
        dev = device or torch.device('cuda', torch.cuda.current_device())
        self.multi = None
        if share_operator and 1 < len(systems) <= MultiRhsCG.MAX_K and \
                _same_operator(systems):
            B = np.stack([s['b'] for s in systems], 1)
            self.multi = MultiRhsCG(systems[0], B, jacobi, dev)
            self._multi_args = (systems[0], B, jacobi, dev)
            self.groups = None        # autotuned split of the right-hand sides
            self.K = len(systems)
            self.sizes = np.full(self.K, systems[0]['Av'].shape[0], np.int64)
            self.offsets = np.concatenate([[0], np.cumsum(self.sizes)])
            self.R = None
            return
        sb = ctypes.c_int64()
        nat.call('esr_bcg_state_bytes', ctypes.byref(sb))
        assert sb.value == _STATE_DTYPE.itemsize, (sb.value, _STATE_DTYPE.itemsize)
        self.K = len(systems)
        self.sizes = np.array([s['Av'].shape[0] for s in systems], np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)])
        N = int(self.offsets[-1])
        nnz = np.array([s['row'].shape[0] for s in systems], np.int64)
        if N >= 2 ** 31 or nnz.sum() >= 2 ** 31:
            raise ValueError('batch too large for int32 indices')
        rowptr = np.zeros(N + 1, np.int64)
        cols, Ae, Av, b = [], [], [], []
        for k, s in enumerate(systems):
            o = self.offsets[k]
            cnt = np.bincount(s['row'], minlength=s['Av'].shape[0])
            rowptr[o + 1:o + 1 + cnt.shape[0]] = cnt
            if np.any(np.diff(s['row']) < 0):
                raise ValueError('system rows must be sorted')
            cols.append(s['col'] + o)
            Ae.append(s['Ae'])
            Av.append(s['Av'])
            b.append(s['b'])
        rowptr = np.cumsum(rowptr)
        t32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(dev)
        f64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)
        self.rowptr, self.col = t32(rowptr), t32(np.concatenate(cols))
        self.Ae, self.Av, self.b = f64(np.concatenate(Ae)), f64(np.concatenate(Av)), \
            f64(np.concatenate(b))
        self.diag = self.Av if jacobi else None
        self.x = torch.zeros(N, dtype=torch.float64, device=dev)
        self.r, self.z, self.p, self.Ap = (torch.empty_like(self.x) for _ in range(4))
        self.state = torch.zeros(self.K * _STATE_DTYPE.itemsize, dtype=torch.uint8,
                                 device=dev)
        self.active = torch.zeros(1, dtype=torch.int32, device=dev)
        self.dev = dev
        self.N = N
        self.R = None
        self.tiles = {}
        self.graphs = {}

    # -- tiling -------------------------------------------------------------
    def _tiling(self, R):
        """Instance k gets ceil(n_k / (256 R)) workgroups (at most
        _MAX_TILES), which stride over its rows."""
        if R not in self.tiles:
            rows = 256 * R
            inst, r0, r1, first, nt = [], [], [], [], []
            for k in range(self.K):
                o, n = int(self.offsets[k]), int(self.sizes[k])
                first.append(len(inst))
                cnt = min(max((n + rows - 1) // rows, 1), _MAX_TILES)
                nt.append(cnt)
                inst.extend([k] * cnt)
                r0.extend([o] * cnt)
                r1.extend([o + n] * cnt)
            t32 = lambda a: torch.tensor(a, dtype=torch.int32, device=self.dev)
            ntiles = len(inst)
            self.tiles[R] = dict(n=ntiles, inst=t32(inst), row0=t32(r0), row1=t32(r1),
                                 first=t32(first), ntiles=t32(nt),
                                 part=torch.empty(3 * ntiles, dtype=torch.float64,
                                                  device=self.dev))
        return self.tiles[R]

    def _tile_ptrs(self, R):
        t = self._tiling(R)
        return [R, t['n'], t['inst'].data_ptr(), t['row0'].data_ptr(),
                t['row1'].data_ptr(), t['first'].data_ptr(), t['ntiles'].data_ptr()]

    def _op(self):
        return [self.rowptr.data_ptr(), self.col.data_ptr(), self.Ae.data_ptr(),
                self.Av.data_ptr(), None if self.diag is None else self.diag.data_ptr()]

    def _init(self, R, rtol, atol, maxiter):
        st = np.zeros(self.K, _STATE_DTYPE)
        st['rtol'] = rtol
        st['atol'] = atol if atol else 0.0
        st['maxiter'] = -1 if maxiter is None else maxiter
        self.state.copy_(torch.from_numpy(st.view(np.uint8)).to(self.dev))
        ops.fill_(self.active, self.K)
        t = self._tiling(R)
        nat.call('esr_bcg_init', *self._tile_ptrs(R), *self._op(), self.b.data_ptr(),
                 self.x.data_ptr(), self.r.data_ptr(), self.z.data_ptr(),
                 self.p.data_ptr(), self.state.data_ptr(), t['part'].data_ptr(),
                 nat.stream_ptr())

    def _steps(self, R, count, sp):
        t = self._tiling(R)
        nat.call('esr_bcg_steps', *self._tile_ptrs(R), *self._op(), self.x.data_ptr(),
                 self.r.data_ptr(), self.z.data_ptr(), self.p.data_ptr(),
                 self.Ap.data_ptr(), self.state.data_ptr(), t['part'].data_ptr(),
                 self.active.data_ptr(), count, sp)

    def _split(self, S):
        """The right-hand sides in S groups of consecutive columns, one
        shared-operator solver per group (solved one after the other)."""
        system, B, jacobi, dev = self._multi_args
        bounds = np.linspace(0, self.K, S + 1).round().astype(int)
        return [MultiRhsCG(system, B[:, a:b], jacobi, dev)
                for a, b in zip(bounds[:-1], bounds[1:]) if b > a]

    def autotune(self, iters=5):
        """Pick the tile height with the fastest iterations on this batch;
        for right-hand sides of one operator, the number of groups the
        right-hand sides are split into (1, 2 or 4 shared-operator solvers run
        one after the other: fewer columns per row gather, the operator read
        once per group)."""
        if self.multi is not None:
            timings = {}
            best = None
            for S in (1, 2, 4):
                if S > self.K:
                    continue
                groups = [self.multi] if S == 1 else self._split(S)
                timings[S] = sum(g.step_time(iters) for g in groups)
                if best is None or timings[S] < timings[best]:
                    best, self.groups = S, groups
            self.autotune_ms = timings
            return best, timings
        best = None
        timings = {}
        for R in _CANDIDATES:
            ops.fill_(self.x, 0)
            self._init(R, 0.0, None, None)
            self._steps(R, 1, nat.stream_ptr())      # warm-up
            torch.cuda.synchronize()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            self._steps(R, iters, nat.stream_ptr())
            b.record()
            torch.cuda.synchronize()
            timings[R] = a.elapsed_time(b) / iters
            if best is None or timings[R] < timings[best]:
                best = R
        self.R = best
        self.autotune_ms = timings
        ops.fill_(self.x, 0)
        return best, timings

    def solve(self, rtol=1e-8, atol=None, maxiter=None, chunk=32):
        """Solve every instance from x = 0; returns a list of
        {'residual', 'iters'} (cg.py:137 per instance)."""
        if self.multi is not None:
            out = []
            for g in (self.groups or [self.multi]):
                out.extend(g.solve(rtol, atol, maxiter, chunk))
            return out
        R = self.R or 4
        ops.fill_(self.x, 0)
        self._init(R, rtol, atol, maxiter)
        key = (R, chunk)
        if key not in self.graphs:
            self.graphs[key] = _Graph(lambda sp: self._steps(R, chunk, sp))
        g = self.graphs[key]
        while True:
            g.replay()
            if int(self.active.item()) == 0:
                break
        st = self.state.cpu().numpy().view(_STATE_DTYPE)
        return [{'residual': float(st['rnorm'][k]), 'iters': int(st['iters'][k])}
                for k in range(self.K)]

    def solutions(self):
        if self.multi is not None:
            out = []
            for g in (self.groups or [self.multi]):
                out.extend(g.solutions())
            return out
        x = self.x.cpu().numpy()
        return [x[self.offsets[k]:self.offsets[k + 1]] for k in range(self.K)]
