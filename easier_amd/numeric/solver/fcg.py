"""Flexible CG with a learned preconditioner and a Jacobi fallback.

No reference counterpart (the reference's CG, numeric/solver/cg.py, assumes
a fixed preconditioner). Each step:

    q = A p; alpha = r.z / p.q; r_old = r; r -= alpha q; x += alpha p
    z = M(r)            (the learned network; Jacobi r/d after a fallback)
    beta = z.(r - r_old) / (r_old.z_old)     (Polak-Ribiere)
    p = z + beta p

Convergence is declared only on the true float64 residual ||b - A x||,
recomputed with the plain operator when the recursive residual reaches the
tolerance (a miss restarts from the true residual). A network output that is
not a descent direction (z.r <= 0 or not finite) is replaced by the Jacobi
vector r / d for that step, on the device. The solver hands the rest of the
solve to the fused Jacobi-PCG when the network was rejected for most steps
of a chunk or when the residual has not improved for `window` steps.
"""
import ctypes
from typing import Any, Dict, Optional

import numpy as np
import torch
from torch import nn

from ... import _native as nat
from ...core import module as esr
from ...core import ops
from ...core.utils import logger
from ..preconditioner.learned import LearnedPreconditioner, MeshGraph
from .fused import FusedCG, LinsysCSR, make_sell
from .mixed import TrueResidual

_c_i32, _c_i64, _c_vp, _c_dbl = ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_double
nat.register_signatures({
    'esr_fcg_workspace': (_c_i32, [_c_i64, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_i64)]),
    'esr_fcg_op': (_c_i32, [_c_i32, _c_i64] + [_c_vp] * 12 + [_c_i32, _c_vp, _c_vp, _c_vp]),
    'esr_fcg_op_sell': (_c_i32, [_c_i32, _c_i64] + [_c_vp] * 3 + [_c_i32] + [_c_vp] * 10
                        + [_c_i32, _c_vp, _c_vp, _c_vp]),
    'esr_fcg_set': (_c_i32, [_c_vp, _c_dbl, _c_i64, _c_vp]),
    'esr_fcg_read': (_c_i32, [_c_vp, ctypes.POINTER(ctypes.c_double),
                              ctypes.POINTER(_c_i64), _c_vp]),
})

SPMV, UPDATE, DOTS, DIRECTION, RESIDUAL, JACOBI = range(6)


class FcgKernels:
    """Device vectors and kernels of one flexible-CG solve on a CSR Linsys
    (rowptr, col, Ae, Av all on the device, float64).

    Steps with the learned preconditioner run in chunks of `chunk` steps
    captured as one HIP graph (network forward included); the device stops
    stepping once ||r|| <= tol, and the host checks progress once per chunk.
    The Jacobi fallback hands the current iterate to the fused Jacobi-PCG
    kernels (the plain solver), so the rest of the solve runs at its speed."""

    def __init__(self, rowptr, col, Ae, Av, n, device, chunk=16):
        self.rowptr, self.col, self.Ae, self.Av, self.n = rowptr, col, Ae, Av, n
        mk = lambda: torch.zeros(n, dtype=torch.float64, device=device)
        self.r, self.ro, self.p, self.q, self.z = mk(), mk(), mk(), mk(), mk()
        sb, wb = ctypes.c_int64(), ctypes.c_int64()
        nat.call('esr_fcg_workspace', n, ctypes.byref(sb), ctypes.byref(wb))
        self.state = torch.zeros(sb.value, dtype=torch.uint8, device=device)
        self.ws = torch.empty(wb.value, dtype=torch.uint8, device=device)
        nat.call('esr_pcg_workspace', n, ctypes.byref(sb), ctypes.byref(wb))
        self.pcg_state = torch.zeros(sb.value, dtype=torch.uint8, device=device)
        self.pcg_ws = torch.empty(wb.value, dtype=torch.uint8, device=device)
        self.chunk = chunk
        self.graphs = {}
        # the operator's SELL-64 form (csrc/esr_krylov.h) for the products,
        # built on the first solve and refreshed from Ae on every solve
        self.sell, self.want_sell = None, FusedCG.use_sell

    def _refresh_operator(self):
        if self.want_sell:
            self.want_sell = False
            self.sell = make_sell(self)
            self.graphs.clear()
        elif self.sell is not None:
            self.sell.refresh(self.Ae)

    def op(self, code, b=None, x=None, a0=0):
        P = lambda t: None if t is None else t.data_ptr()
        vecs = [P(self.Av), P(b), P(x), P(self.r), P(self.ro), P(self.p), P(self.q),
                P(self.z), a0, P(self.state), P(self.ws), nat.stream_ptr()]
        if self.sell is not None:
            nat.call('esr_fcg_op_sell', code, self.n, *self.sell.arrays(), P(self.Av), *vecs)
        else:
            nat.call('esr_fcg_op', code, self.n, P(self.rowptr), P(self.col), P(self.Ae),
                     P(self.Av), *vecs)

    def read(self):
        sc = (ctypes.c_double * 9)()
        it = (ctypes.c_int64 * 3)()
        nat.call('esr_fcg_read', self.state.data_ptr(), sc, it, nat.stream_ptr())
        keys = ('rz', 'pq', 'alpha', 'rr', 'zr', 'zro', 'beta', 'rr_true', 'bb')
        out = dict(zip(keys, sc))
        out.update(it=it[0], done=it[1], nbad=it[2])
        return out

    def _step(self, M, x, restart):
        """One flexible-CG step: z = M(r), beta, p, q = A p, alpha, x, r. A
        z that is not a descent direction is replaced by r / d (Jacobi) for
        this step, on the device."""
        M.apply(self.r, self.z)
        self.op(DOTS, a0=1 if restart else 0)
        self.op(DIRECTION, a0=1 if restart else 0)
        self.op(SPMV)
        self.op(UPDATE, x=x)

    def _chunk_graph(self, M, x):
        # the graph holds the network's buffers of its current binding
        key = (id(M), M.binding, x.data_ptr())
        if key not in self.graphs:
            from .fused import _Graph

            def launch(sp):
                for _ in range(self.chunk):
                    self._step(M, x, False)
            self.graphs[key] = (_Graph(launch), M)
        return self.graphs[key][0]

    def _jacobi_pcg(self, b, x, tol, maxiter):
        """Fused Jacobi-PCG from the current x to ||r|| <= tol; returns
        (steps, recursive residual)."""
        from .fused import _Graph
        P = lambda t: t.data_ptr()
        csr = [P(self.rowptr), P(self.col), P(self.Ae), P(self.Av), P(self.Av)]
        nat.call('esr_pcg_init', nat.F64, self.n, *csr, P(b), P(x), P(self.r), P(self.z),
                 P(self.p), P(self.pcg_state), P(self.pcg_ws), 0.0, float(tol),
                 -1 if maxiter is None else int(maxiter), nat.stream_ptr())
        if 'pcg' not in self.graphs:
            def launch(sp):
                vecs = [P(x), P(self.r), P(self.z), P(self.p), P(self.q), P(self.pcg_state),
                        P(self.pcg_ws), 32, sp]
                if self.sell is not None:
                    nat.call('esr_pcg_steps_sell', nat.F64, self.n, *self.sell.arrays(),
                             P(self.Av), P(self.Av), *vecs)
                else:
                    nat.call('esr_pcg_steps', nat.F64, self.n, *csr, *vecs)
            self.graphs['pcg'] = (_Graph(launch), x.data_ptr())
        g, xp = self.graphs['pcg']
        assert xp == x.data_ptr()
        sc = (ctypes.c_double * 6)()
        it = (ctypes.c_int64 * 3)()
        while True:
            g.replay()
            nat.call('esr_pcg_read', self.pcg_state.data_ptr(), sc, it, nat.stream_ptr())
            if it[1]:
                return int(it[0]), float(sc[0])

    def solve(self, b, x, M, rtol=1e-8, atol=None, maxiter=None, window=200,
              debug_iter=None):
        self._refresh_operator()
        self.op(RESIDUAL, b=b, x=x)
        s = self.read()
        bnorm = float(np.sqrt(s['bb']))
        tol = max(rtol * bnorm, atol) if atol else rtol * bnorm
        rnorm = float(np.sqrt(s['rr_true']))
        fallback_at, why, rejected = None, None, 0
        it = 0
        if M is not None and not rnorm <= tol:
            best, best_it = rnorm, 0
            restart = True
            while maxiter is None or it < maxiter:
                left = -1 if maxiter is None else maxiter - it
                nat.call('esr_fcg_set', self.state.data_ptr(), tol * tol, left,
                         nat.stream_ptr())
                if restart:
                    self._step(M, x, True)
                    restart = False
                self._chunk_graph(M, x).replay()
                s = self.read()
                it += s['it']
                rnorm = float(np.sqrt(s['rr']))
                if debug_iter:
                    logger.info(f'FlexibleCG residual {rnorm} at step {it}')
                if s['done'] and rnorm <= tol:
                    self.op(RESIDUAL, b=b, x=x)          # judge on the true residual
                    rnorm = float(np.sqrt(self.read()['rr_true']))
                    if rnorm <= tol:
                        break
                    restart = True                       # drifted: restart from it
                    continue
                if rnorm < best * 0.999:
                    best, best_it = rnorm, it
                rejected += s['nbad']
                if 2 * s['nbad'] > s['it']:
                    why = 'non-descent directions'
                elif it - best_it > window:
                    why = 'stagnation'
                if why:
                    fallback_at = it
                    logger.info(f'FlexibleCG: falling back to Jacobi-PCG at step {it} ({why})')
                    break
        if M is None or why:
            # the plain solver from the current iterate, judged on the true residual
            while maxiter is None or it < maxiter:
                k, _ = self._jacobi_pcg(b, x, tol, None if maxiter is None else maxiter - it)
                it += max(k, 1)
                self.op(RESIDUAL, b=b, x=x)
                rnorm = float(np.sqrt(self.read()['rr_true']))
                if rnorm <= tol:
                    break
        return {'residual': rnorm, 'iters': it, 'fallback_at': fallback_at,
                'rejected_steps': rejected,
                'bnorm': bnorm}


class FlexibleCG(esr.Module):
    """Flexible CG on a compiled Linsys with an optional learned or multigrid
    preconditioner. `mesh_features` = dict(points (n, 2), logk (n,), n_grid)
    in the user's vertex order: the network's inputs (edge vectors in units
    of 1/n_grid) and the grid of geometric coarsening, renumbered with the
    program after esr.compile(). A multigrid preconditioner without a
    network needs none of them: without a grid it coarsens by aggregation."""

    def __init__(self, A: nn.Module, b: esr.Tensor, x: esr.Tensor,
                 M: Optional[LearnedPreconditioner] = None,
                 mesh_features: Optional[dict] = None) -> None:
        super().__init__()
        self.A = A
        self.b = b
        self.x = x
        self.easier_M = M
        self.easier_features = mesh_features
        self.r = esr.Tensor(esr.zeros_like(x), mode='partition')
        self.rnorm = esr.Tensor(torch.tensor([0.0], dtype=x.dtype), mode='replicate')
        self.residual = TrueResidual(A, b, x, self.r, self.rnorm)
        self.easier_kernels = None

    def easier_after_compile(self, prog):
        why = LinsysCSR.applicable(self.A, self.x, prog)
        if why or self.x.dtype != torch.float64:
            raise RuntimeError(f'FlexibleCG needs a float64 Linsys ({why})')
        op = LinsysCSR(self.A)
        # steps per captured chunk: a strong preconditioner converges in few
        # steps, and steps after convergence still run its cycle
        self.easier_kernels = FcgKernels(op.rowptr, op.col, op.Ae, op.Av, op.n,
                                         self.x.device,
                                         chunk=getattr(self.easier_M, 'fcg_chunk', 16))
        if self.easier_M is not None:
            from ..preconditioner.multigrid import MultigridPreconditioner, grid_positions
            M = self.easier_M
            f = self.easier_features
            mg = isinstance(M, MultigridPreconditioner)
            learned = M.smoother if mg else M
            if f is None and learned is not None:
                raise RuntimeError('a learned preconditioner needs mesh_features')
            points = logk = n_grid = None
            if f is not None:
                o = self.x.easier_order
                o = np.arange(op.n) if o is None else o.numpy()
                points = np.asarray(f['points'])[o]
                logk = None if f.get('logk') is None else np.asarray(f['logk']).reshape(-1)[o]
                n_grid = f.get('n_grid')
            graph = None
            if learned is not None:
                rowptr = self.A.reducer.easier_host_rowptr
                row = np.repeat(np.arange(op.n), np.diff(rowptr))
                graph = MeshGraph(row, self.A.selector.easier_host_idx,
                                  op.Av.cpu().numpy(), op.Ae.cpu().numpy(), points, logk,
                                  n_grid, self.x.device)
            if mg:
                # geometric coarsening needs the vertices on the n_grid lattice
                fpos, grid = None, None
                if points is not None and n_grid is not None and M.coarsening != 'aggregation':
                    try:
                        fpos, grid = grid_positions(points, n_grid), n_grid
                    except ValueError:
                        if M.coarsening == 'geometric':
                            raise
                M.setup(op.rowptr, op.col, op.Ae, op.Av, grid, fpos, logk, graph)
            else:
                M.bind(graph)

    def solve(self, rtol=1e-8, atol=None, maxiter=None, window=200,
              debug_iter=None) -> Dict[str, Any]:
        if self.easier_kernels is None:
            raise RuntimeError('run esr.compile() first')
        return self.easier_kernels.solve(self.b.data, self.x.data, self.easier_M, rtol,
                                         atol, maxiter, window, debug_iter)
