"""Flexible CG with a learned preconditioner and a Jacobi fallback.

No reference counterpart (the reference's CG, numeric/solver/cg.py, assumes
a fixed preconditioner). Each step:

    q = A p; alpha = r.z / p.q; r_old = r; r -= alpha q; x += alpha p
    z = M(r)            (the learned network; Jacobi r/d after a fallback)
    beta = z.(r - r_old) / (r_old.z_old)     (Polak-Ribiere)
    p = z + beta p

Convergence is declared only on the true float64 residual ||b - A x||,
recomputed with the plain operator when the recursive residual reaches the
tolerance (a miss restarts from the true residual). The solver falls back to
Jacobi-PCG for the rest of the solve when the network returns a non-descent
direction (z.r <= 0 or not finite) or when the residual has not improved for
`window` steps.
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
from .fused import LinsysCSR
from .mixed import TrueResidual

_c_i32, _c_i64, _c_vp = ctypes.c_int, ctypes.c_int64, ctypes.c_void_p
nat.register_signatures({
    'esr_fcg_workspace': (_c_i32, [_c_i64, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_i64)]),
    'esr_fcg_op': (_c_i32, [_c_i32, _c_i64] + [_c_vp] * 12 + [_c_i32, _c_vp, _c_vp, _c_vp]),
    'esr_fcg_read': (_c_i32, [_c_vp, ctypes.POINTER(ctypes.c_double), _c_vp]),
})

SPMV, UPDATE, DOTS, DIRECTION, RESIDUAL, JACOBI = range(6)


class FcgKernels:
    """Device vectors and kernels of one flexible-CG solve on a CSR Linsys
    (rowptr, col, Ae, Av all on the device, float64)."""

    def __init__(self, rowptr, col, Ae, Av, n, device):
        self.rowptr, self.col, self.Ae, self.Av, self.n = rowptr, col, Ae, Av, n
        mk = lambda: torch.zeros(n, dtype=torch.float64, device=device)
        self.r, self.ro, self.p, self.q, self.z = mk(), mk(), mk(), mk(), mk()
        sb, wb = ctypes.c_int64(), ctypes.c_int64()
        nat.call('esr_fcg_workspace', n, ctypes.byref(sb), ctypes.byref(wb))
        self.state = torch.zeros(sb.value, dtype=torch.uint8, device=device)
        self.ws = torch.empty(wb.value, dtype=torch.uint8, device=device)

    def op(self, code, b=None, x=None, a0=0):
        P = lambda t: None if t is None else t.data_ptr()
        nat.call('esr_fcg_op', code, self.n, P(self.rowptr), P(self.col), P(self.Ae),
                 P(self.Av), P(self.Av), P(b), P(x), P(self.r), P(self.ro), P(self.p),
                 P(self.q), P(self.z), a0, P(self.state), P(self.ws), nat.stream_ptr())

    def read(self):
        sc = (ctypes.c_double * 9)()
        nat.call('esr_fcg_read', self.state.data_ptr(), sc, nat.stream_ptr())
        keys = ('rz', 'pq', 'alpha', 'rr', 'zr', 'zro', 'beta', 'rr_true', 'bb')
        return dict(zip(keys, sc))

    def precondition(self, M: Optional[LearnedPreconditioner], use_jacobi: bool):
        if M is None or use_jacobi:
            self.op(JACOBI)
        else:
            M.apply(self.r, self.z)

    def solve(self, b, x, M, rtol=1e-8, atol=None, maxiter=None, window=50,
              debug_iter=None):
        self.op(RESIDUAL, b=b, x=x)
        s = self.read()
        bnorm = float(np.sqrt(s['bb']))
        tol = max(rtol * bnorm, atol) if atol else rtol * bnorm
        rnorm = float(np.sqrt(s['rr_true']))
        use_jacobi = M is None
        fallback_at = None
        best, best_it = rnorm, 0
        it = 0
        restart = True
        while True:
            if rnorm == rnorm and rnorm <= tol:
                self.op(RESIDUAL, b=b, x=x)
                s = self.read()
                rnorm = float(np.sqrt(s['rr_true']))
                if rnorm <= tol:
                    break
                restart = True          # recursive residual drifted: restart
            if maxiter is not None and it >= maxiter:
                break
            self.precondition(M, use_jacobi)
            self.op(DOTS, a0=1 if restart else 0)
            s = self.read()
            bad = not (s['zr'] > 0.0) or not np.isfinite(s['zr'])
            stalled = it - best_it > window
            if not use_jacobi and (bad or stalled):
                use_jacobi = True
                fallback_at = it
                logger.info(f'FlexibleCG: falling back to Jacobi at step {it} '
                            f'({"non-descent direction" if bad else "stagnation"})')
                self.op(JACOBI)
                self.op(DOTS, a0=1)
                restart = True
            if restart:
                ops.copy_(self.p, self.z)
                restart = False
            else:
                self.op(DIRECTION)
            self.op(SPMV)
            self.op(UPDATE, x=x)
            it += 1
            s = self.read()
            rnorm = float(np.sqrt(s['rr']))
            if rnorm < best * 0.999:
                best, best_it = rnorm, it
            if debug_iter and it % debug_iter == 0:
                logger.info(f'FlexibleCG residual {rnorm} at step {it}')
        return {'residual': rnorm, 'iters': it, 'fallback_at': fallback_at,
                'bnorm': bnorm}


class FlexibleCG(esr.Module):
    """Flexible CG on a compiled Linsys with an optional learned
    preconditioner. `mesh_features` = dict(points (n, 2), logk (n,), n_grid)
    in the user's vertex order: the network's inputs, renumbered with the
    program after esr.compile()."""

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
        self.easier_kernels = FcgKernels(op.rowptr, op.col, op.Ae, op.Av, op.n,
                                         self.x.device)
        if self.easier_M is not None:
            f = self.easier_features
            if f is None:
                raise RuntimeError('a learned preconditioner needs mesh_features')
            o = self.x.easier_order
            o = np.arange(op.n) if o is None else o.numpy()
            rowptr = self.A.reducer.easier_host_rowptr
            row = np.repeat(np.arange(op.n), np.diff(rowptr))
            graph = MeshGraph(row, self.A.selector.easier_host_idx,
                              op.Av.cpu().numpy(), op.Ae.cpu().numpy(),
                              np.asarray(f['points'])[o], np.asarray(f['logk']).reshape(-1)[o],
                              f['n_grid'], self.x.device)
            self.easier_M.bind(graph)

    def solve(self, rtol=1e-8, atol=None, maxiter=None, window=50,
              debug_iter=None) -> Dict[str, Any]:
        if self.easier_kernels is None:
            raise RuntimeError('run esr.compile() first')
        return self.easier_kernels.solve(self.b.data, self.x.data, self.easier_M, rtol,
                                         atol, maxiter, window, debug_iter)
