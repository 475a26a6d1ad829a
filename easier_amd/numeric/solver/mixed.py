"""Mixed-precision conjugate gradients by iterative refinement.

No reference counterpart (the reference's CG, numeric/solver/cg.py, runs in
the tensors' dtype). The outer loop is float64:

    r = b - A x                      (float64 operator, ||r|| in float64)
    stop when ||r|| <= max(rtol ||b||, atol)
    solve A d = r to `inner_rtol` with Jacobi-PCG in float32
        (SpMV, axpy, vectors and operator in float32; dot products in float64)
    x += d                            (float64)

so convergence is always judged on the true float64 residual. The inner
solve is the fused CG iteration (solver/fused.py) on float32 copies of the
operator, which halves the bytes an iteration moves.
"""
import ctypes
from typing import Any, Dict, Optional

import torch
from torch import nn

from ... import _native as nat
from ...core import module as esr
from ...core import ops
from ...core.fuser import BoundTensors
from ...core.utils import logger
from .fused import FusedCG, LinsysCSR, _Graph, _precond_diag, make_sell

_c_i32, _c_i64, _c_vp = ctypes.c_int, ctypes.c_int64, ctypes.c_void_p
nat.register_signatures({
    'esr_ir_workspace': (_c_i32, [_c_i64, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_i64)]),
    'esr_ir_residual': (_c_i32, [_c_i64] + [_c_vp] * 9 + [_c_i32, _c_vp]),
    'esr_ir_update': (_c_i32, [_c_i64, _c_vp, _c_vp, _c_vp]),
    'esr_ir_read': (_c_i32, [_c_vp, ctypes.POINTER(ctypes.c_double), _c_vp]),
    'esr_pcg_rearm': (_c_i32, [_c_i32, _c_i64] + [_c_vp] * 6 + [ctypes.c_double, _c_vp]),
})


class TrueResidual(esr.Module):
    """r = b - A x in the tensors' dtype (the role of gmres.py's UpdateRnorm);
    traced by esr.compile, it also ties A's index spaces to x."""

    def __init__(self, A, b, x, r, rnorm):
        super().__init__()
        self.A = A
        self.b = b
        self.x = x
        self.r = r
        self.rnorm = rnorm

    def forward(self):
        r = self.b - self.A(self.x)
        self.r[:] = r
        self.rnorm[:] = esr.norm(r).norm()


class MixedPrecisionCG(esr.Module):
    """CG by float64 iterative refinement around a float32 inner CG.

    Same constructor as CG (A: a Linsys in float64, b, x: float64 partition
    tensors, M: None or preconditioner.Jacobi)."""

    def __init__(self, A: nn.Module, b: esr.Tensor, x: esr.Tensor,
                 M: Optional[nn.Module] = None) -> None:
        super().__init__()
        self.A = A
        self.b = b
        self.x = x
        self.M = M
        self.r = esr.Tensor(esr.zeros_like(x), mode='partition')
        self.rnorm = esr.Tensor(torch.tensor([0.0], dtype=x.dtype),
                                mode='replicate')
        self.residual = TrueResidual(A, b, x, self.r, self.rnorm)
        self.easier_state = None

    def easier_after_compile(self, prog):
        why = LinsysCSR.applicable(self.A, self.x, prog)
        ok, diag = _precond_diag(self.M, self.x, prog)
        if why or not ok or self.x.dtype != torch.float64:
            raise RuntimeError('MixedPrecisionCG needs a float64 Linsys and an '
                               f'identity/Jacobi preconditioner ({why})')
        op = LinsysCSR(self.A)
        dev = self.x.device
        n = op.n
        f32 = lambda t: None if t is None else t.to(torch.float32)
        self.easier_state = s = dict(op=op, n=n, diag64=diag)
        s['bound'] = BoundTensors(op.sources + [self.b, self.x])
        s['Ae32'], s['Av32'] = f32(op.Ae), f32(op.Av)
        # M(r) = r / A.Av (the usual Jacobi form): one float32 copy serves both
        same = diag is not None and diag.data_ptr() == op.Av.data_ptr()
        s['diag32'] = s['Av32'] if same else f32(diag)
        s['want_sell'] = FusedCG.use_sell      # built on the first solve
        s['sell'] = None
        for k in ('r32', 'd32', 'ri', 'z', 'p', 'Ap'):
            s[k] = torch.zeros(n, dtype=torch.float32, device=dev)
        sb, wb = ctypes.c_int64(), ctypes.c_int64()
        nat.call('esr_pcg_workspace', n, ctypes.byref(sb), ctypes.byref(wb))
        s['pcg_state'] = torch.empty(sb.value, dtype=torch.uint8, device=dev)
        s['pcg_ws'] = torch.empty(wb.value, dtype=torch.uint8, device=dev)
        nat.call('esr_ir_workspace', n, ctypes.byref(sb), ctypes.byref(wb))
        s['ir_state'] = torch.empty(sb.value, dtype=torch.uint8, device=dev)
        s['ir_ws'] = torch.empty(wb.value, dtype=torch.uint8, device=dev)
        s['graph'] = None

    def _inner_arrays(self):
        s = self.easier_state
        op = s['op']
        return [op.rowptr.data_ptr(), op.col.data_ptr(), s['Ae32'].data_ptr(),
                s['Av32'].data_ptr(),
                None if s['diag32'] is None else s['diag32'].data_ptr()]

    def _inner_steps(self, count, sp):
        s = self.easier_state
        vecs = [s['d32'].data_ptr(), s['ri'].data_ptr(), s['z'].data_ptr(),
                s['p'].data_ptr(), s['Ap'].data_ptr(), s['pcg_state'].data_ptr(),
                s['pcg_ws'].data_ptr(), count, sp]
        if s['sell'] is not None:
            nat.call('esr_pcg_steps_sell', nat.F32, s['n'], *s['sell'].arrays(),
                     s['Av32'].data_ptr(),
                     None if s['diag32'] is None else s['diag32'].data_ptr(), *vecs)
        else:
            nat.call('esr_pcg_steps', nat.F32, s['n'], *self._inner_arrays(), *vecs)

    def _refresh_operator(self):
        """The float32 copies of the operator (and the Jacobi diagonal) from
        the float64 tensors as they are now, so values changed in place
        between solves are seen."""
        s = self.easier_state
        op = s['op']
        for src, dst in ((op.Ae, s['Ae32']), (op.Av, s['Av32']), (s['diag64'], s['diag32'])):
            if src is not None and not (dst is s['Av32'] and src is not op.Av):
                ops.ewise('copy', src, out=dst)
        if s['want_sell']:
            s['want_sell'] = False
            s['sell'] = make_sell(op, s['Ae32'])
            s['graph'] = None
        elif s['sell'] is not None:
            s['sell'].refresh(s['Ae32'])

    def _residual(self, first):
        s = self.easier_state
        op = s['op']
        nat.call('esr_ir_residual', s['n'], op.rowptr.data_ptr(), op.col.data_ptr(),
                 op.Ae.data_ptr(), op.Av.data_ptr(), self.b.data.data_ptr(),
                 self.x.data.data_ptr(), s['r32'].data_ptr(), s['ir_state'].data_ptr(),
                 s['ir_ws'].data_ptr(), 1 if first else 0, nat.stream_ptr())
        sc = (ctypes.c_double * 2)()
        nat.call('esr_ir_read', s['ir_state'].data_ptr(), sc, nat.stream_ptr())
        return sc[0], sc[1]

    def solve(self, rtol: float = 1e-10, atol: Optional[float] = None,
              maxiter: Optional[int] = None, inner_rtol: float = 1e-4,
              max_outer: int = 200, chunk: int = 32, reliable: bool = True,
              delta: float = 0.5,
              debug_iter: Optional[int] = None) -> Dict[str, Any]:
        """Returns {'residual': true float64 ||b - A x||, 'iters': float32
        steps, 'outer': float64 residual updates, 'bnorm'}.

        reliable=True (default): one float32 CG whose correction is folded
        into the float64 x, and whose residual is replaced by the true
        float64 one, each time the recursive residual has dropped by `delta`
        since the last update; the search direction is kept across updates
        (reliable updates, Sleijpen and van der Vorst 1996), so the step count
        stays close to float64 CG. maxiter caps the total float32 steps.
        (Variable-coefficient Poisson, 4.2 M vertices, to 1e-10: delta 0.1 /
        0.3 / 0.5 / 0.7 took 11,068 / 10,659 / 10,093 / 10,073 float32 steps,
        1.20 / 1.16 / 1.10 / 1.11 s; float64 CG 9,083 steps, 1.45 s.)
        reliable=False: classic iterative refinement, a fresh float32 solve
        to `inner_rtol` per outer step (maxiter caps each)."""
        if self.easier_state is None:
            raise RuntimeError('run esr.compile() first')
        s = self.easier_state
        s['bound'].check()
        self._refresh_operator()
        if s['graph'] is None or s.get('chunk') != chunk:
            s['graph'] = _Graph(lambda sp: self._inner_steps(chunk, sp))
            s['chunk'] = chunk
        rnorm, bnorm = self._residual(True)
        tol = max(rtol * bnorm, atol) if atol else rtol * bnorm
        total, outer = 0, 0
        sc = (ctypes.c_double * 6)()
        it = (ctypes.c_int64 * 3)()

        def run_inner():
            while True:
                s['graph'].replay()
                nat.call('esr_pcg_read', s['pcg_state'].data_ptr(), sc, it,
                         nat.stream_ptr())
                if it[1]:
                    return

        started = False
        while not (rnorm == rnorm and rnorm <= tol) and outer < max_outer:
            outer += 1
            if reliable and started:
                if maxiter is not None and it[0] >= maxiter:
                    break
                # the inner residual is the true one again; keep p
                ops.copy_(s['ri'], s['r32'])
                nat.call('esr_pcg_rearm', nat.F32, s['n'],
                         None if s['diag32'] is None else s['diag32'].data_ptr(),
                         s['ri'].data_ptr(), s['z'].data_ptr(), s['p'].data_ptr(),
                         s['pcg_state'].data_ptr(), s['pcg_ws'].data_ptr(),
                         max(delta * rnorm, tol), nat.stream_ptr())
            else:
                # float32 Jacobi-PCG on A d = r from d = 0
                inner_tol = max(delta * rnorm, tol) if reliable else inner_rtol * rnorm
                nat.call('esr_pcg_init', nat.F32, s['n'], *self._inner_arrays(),
                         s['r32'].data_ptr(), s['d32'].data_ptr(), s['ri'].data_ptr(),
                         s['z'].data_ptr(), s['p'].data_ptr(), s['pcg_state'].data_ptr(),
                         s['pcg_ws'].data_ptr(), 0.0, float(inner_tol),
                         -1 if maxiter is None else int(maxiter), nat.stream_ptr())
                started = True
            run_inner()
            if not reliable:
                total += it[0] + 1
            nat.call('esr_ir_update', s['n'], self.x.data.data_ptr(),
                     s['d32'].data_ptr(), nat.stream_ptr())
            rnorm, _ = self._residual(False)
            if debug_iter is not None:
                logger.info(f'MixedPrecisionCG update {outer}: float32 steps '
                            f'{it[0] + 1}, true residual {rnorm}')
        if reliable:
            total = it[0] + 1 if started else 0
        return {'residual': rnorm, 'iters': total, 'outer': outer,
                'bnorm': bnorm}
