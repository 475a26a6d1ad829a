"""Fused device iterations for CG and GMRES on a compiled Linsys.

After esr.compile(..., 'hip'), CG/GMRES check whether their traced forwards
apply an operator of the Linsys form (reducer(selector(x) * Ae) + x * Av,
recognised in the graph: solver/recognize.py) whose Reducer fixes its edge
order (so the edge list is a row-sorted CSR) and a preconditioner that
computes M(r) = r or M(r) = r / d. If so, solve() runs:

CG: one init kernel, then chunks of iterations, each iteration three kernels
    (SpMV + p.Ap; x/r/z update + r.r, r.z; loop test + p update) with every
    scalar (alpha, beta, rz, ||r||, iters, done) on the device. A chunk is
    captured once as a HIP graph and replayed; the host reads the state once
    per chunk. Iterations after the exit test are no-ops, so `iters`,
    `residual` and x are those of cg.py:106-137.
GMRES: per restart cycle one residual kernel, the whole Arnoldi process as one
    captured graph (per step: SpMV, then w orthogonalised by classical
    Gram-Schmidt applied twice -- three reductions, one kernel each, where the
    reference's modified Gram-Schmidt takes j + 2), one host read of H, the
    small SVD least-squares solve on the host, one update kernel.
"""
import ctypes
from typing import Optional

import numpy as np
import torch

from ... import _native as nat
from ...core import module as esr
from ...core import ops
from ...core.fuser import BoundTensors
from ...core.utils import logger
from ..linsys import Linsys
from ..preconditioner.jacobi import Jacobi

_c_i32, _c_i64, _c_dbl, _c_vp = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
_P_i64 = ctypes.POINTER(ctypes.c_int64)
_P_dbl = ctypes.POINTER(ctypes.c_double)

nat.register_signatures({
    'esr_pcg_workspace': (_c_i32, [_c_i64, _P_i64, _P_i64]),
    'esr_pcg_init': (_c_i32, [_c_i32, _c_i64] + [_c_vp] * 12 + [_c_dbl, _c_dbl, _c_i64, _c_vp]),
    'esr_pcg_steps': (_c_i32, [_c_i32, _c_i64] + [_c_vp] * 12 + [_c_i64, _c_vp]),
    'esr_pcg_read': (_c_i32, [_c_vp, _P_dbl, _P_i64, _c_vp]),
    'esr_sell_plan': (_c_i32, [_c_i64] + [_c_vp] * 4 + [_P_i64, ctypes.POINTER(_c_i32), _c_vp]),
    'esr_sell_fill': (_c_i32, [_c_i32, _c_i64] + [_c_vp] * 4 + [_c_i32] + [_c_vp] * 4),
    'esr_sell_values': (_c_i32, [_c_i32, _c_i64] + [_c_vp] * 5),
    'esr_pcg_steps_sell': (_c_i32, [_c_i32, _c_i64] + [_c_vp] * 3 + [_c_i32] + [_c_vp] * 10
                           + [_c_i64, _c_vp]),
    'esr_gmres_workspace': (_c_i32, [_c_i64, _c_i64, _P_i64, _P_i64]),
    'esr_gmres_begin': (_c_i32, [_c_i32, _c_i64, _c_vp, _c_vp, _c_vp, _c_vp]),
    'esr_gmres_residual': (_c_i32, [_c_i32, _c_i64] + [_c_vp] * 9 + [_c_vp]),
    'esr_gmres_cycle': (_c_i32, [_c_i32, _c_i64, _c_i64] + [_c_vp] * 11 + [_c_vp]),
    'esr_gmres_update_x': (_c_i32, [_c_i32, _c_i64, _c_i64, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp]),
    'esr_gmres_read': (_c_i32, [_c_vp, _c_i64, _P_dbl, _P_i64, _P_dbl, _c_vp]),
    'esr_gmres_loop_arm': (_c_i32, [_c_vp, _c_dbl, _c_dbl, _c_i64, _c_vp]),
    'esr_gmres_loop_cycles': (_c_i32, [_c_i32, _c_i64, _c_i64] + [_c_vp] * 13 + [_c_i64, _c_vp]),
    'esr_gmres_loop_read': (_c_i32, [_c_vp, _P_dbl, _P_i64, _c_vp]),
})


def _parts(A):
    """{'selector', 'reducer', 'Ae', 'Av'} of a Linsys, or the dict itself
    (an operator recognised in a traced graph, solver/recognize.py)."""
    if isinstance(A, dict):
        return A
    if type(A) is Linsys:
        return dict(selector=A.selector, reducer=A.reducer, Ae=A.Ae, Av=A.Av)
    return None


class LinsysCSR:
    """The device CSR view of a compiled Linsys (or of an operator of the
    same form recognised in the traced forward): rowptr, col (the Selector's
    renumbered index in row-sorted edge order), Ae, Av."""

    def __init__(self, A):
        P = _parts(A)
        self.rowptr = P['reducer'].easier_rowptr
        self.col = P['selector'].idx
        self.Ae = P['Ae'].data
        self.Av = P['Av'].data
        self.n = P['reducer'].n
        self.dtype = self.Av.dtype
        self.sources = [P['Ae'], P['Av']]

    @staticmethod
    def applicable(A, x: esr.Tensor, prog) -> Optional[str]:
        """None if the fused path applies, else the reason it does not."""
        P = _parts(A)
        if P is None:
            return 'operator is not reducer(selector(x) * Ae) + x * Av'
        sel, red = P['selector'], P['reducer']
        if not isinstance(sel, esr.Selector) or not isinstance(red, esr.Reducer):
            return 'Linsys without an esr Selector/Reducer'
        if red.reduce != 'sum':
            return 'Reducer is not a sum'
        if red.easier_perm is not None:
            return 'the Reducer does not fix the edge order'
        Av, Ae = P['Av'], P['Ae']
        for t in (Av, Ae, x):
            if not isinstance(t, esr.Tensor) or not t.is_partition:
                return 'Av/Ae/x are not partition tensors'
        if Av.data.dim() != 1 or Ae.data.dim() != 1 or x.data.dim() != 1:
            return 'only vectors (one value per vertex/edge) are fused'
        dts = {Av.dtype, Ae.dtype, x.dtype}
        if len(dts) != 1 or Av.dtype not in (torch.float32, torch.float64):
            return 'mixed or non-float dtypes'
        if Ae.data.shape[0] != sel.idx.shape[0] or red.n != x.data.shape[0]:
            return 'shape mismatch'
        sa = prog.spaces
        if sa.root(sa.tensor_var[Av]) != sa.root(sa.tensor_var[x]):
            return 'Av and x are different spaces'
        sa = prog.spaces
        sv = sa.root(sa.sel_vars[sel][0])
        rv = sa.root(sa.red_vars[red][1])
        xv = sa.root(sa.tensor_var[x])
        if not (sv == rv == xv):
            return 'selector source, reducer target and x are different spaces'
        return None


class LinsysSELL:
    """The operator of a LinsysCSR in sliced ELLPACK with slices of one
    wavefront (SELL-64, csrc/esr_krylov.h): soff (slice offsets), cnt (row
    lengths, uint8), colv (int16 offsets from the row when every one fits,
    else int32 indices), val (values in `values`' dtype; the CSR's Ae by
    default). The structure is fixed by the compiled Selector and Reducer;
    refresh() copies new operator values in (a solve calls it first, so
    values changed in place between solves are seen). Raises EsrNativeError
    when a row has more than 255 entries (the CSR path is then used)."""

    def __init__(self, csr: 'LinsysCSR', values: Optional[torch.Tensor] = None):
        vals = csr.Ae if values is None else values
        n = csr.n
        dev = csr.Av.device
        self.n, self.rowptr, self.dt = n, csr.rowptr, nat.dtype_code(vals.dtype)
        self.soff = torch.empty((n + 63) // 64 + 1, dtype=torch.int64, device=dev)
        flags = torch.empty(2, dtype=torch.int32, device=dev)
        slots, idx16 = ctypes.c_int64(), ctypes.c_int32()
        nat.call('esr_sell_plan', n, csr.rowptr.data_ptr(), csr.col.data_ptr(),
                 self.soff.data_ptr(), flags.data_ptr(), ctypes.byref(slots),
                 ctypes.byref(idx16), nat.stream_ptr())
        self.idx16 = int(idx16.value)
        self.slots = int(slots.value)
        m = max(self.slots, 1)
        self.cnt = torch.empty(n, dtype=torch.uint8, device=dev)
        self.colv = torch.empty(m, dtype=torch.int16 if self.idx16 else torch.int32,
                                device=dev)
        self.val = torch.empty(m, dtype=vals.dtype, device=dev)
        nat.call('esr_sell_fill', self.dt, n, csr.rowptr.data_ptr(), csr.col.data_ptr(),
                 vals.data_ptr(), self.soff.data_ptr(), self.idx16, self.cnt.data_ptr(),
                 self.colv.data_ptr(), self.val.data_ptr(), nat.stream_ptr())

    def refresh(self, values: torch.Tensor):
        if nat.dtype_code(values.dtype) != self.dt:
            raise TypeError('SELL values keep their dtype')
        nat.call('esr_sell_values', self.dt, self.n, self.rowptr.data_ptr(),
                 values.data_ptr(), self.soff.data_ptr(), self.val.data_ptr(),
                 nat.stream_ptr())

    def arrays(self):
        return [self.soff.data_ptr(), self.cnt.data_ptr(), self.colv.data_ptr(),
                self.idx16, self.val.data_ptr()]


def make_sell(csr: 'LinsysCSR', values: Optional[torch.Tensor] = None):
    """LinsysSELL of the operator, or None when the SELL form does not apply
    (a row of more than 255 entries)."""
    try:
        return LinsysSELL(csr, values)
    except nat.EsrNativeError as e:
        logger.info(f'Linsys kept in CSR form ({e})')
        return None


def _diag_ok(d, x: esr.Tensor, prog) -> bool:
    """A Jacobi diagonal the fused kernels take: a partition vector in x's
    space and dtype."""
    sa = prog.spaces
    return isinstance(d, esr.Tensor) and d.is_partition and d.data.dim() == 1 and \
        d.dtype == x.dtype and sa.root(sa.tensor_var[d]) == sa.root(sa.tensor_var[x])


def _precond_diag(M, x: esr.Tensor, prog):
    """(ok, diag tensor or None) for identity / Jacobi preconditioners."""
    if M is None or getattr(M, '__name__', '') == '_identity':
        return True, None
    if type(M) is Jacobi and isinstance(M.diag, esr.Tensor) and M.diag.is_partition:
        sa = prog.spaces
        if sa.root(sa.tensor_var[M.diag]) == sa.root(sa.tensor_var[x]) and \
                M.diag.dtype == x.dtype and M.diag.data.dim() == 1:
            return True, M.diag.data
    return False, None


def _capture(launch):
    """Capture `launch(stream_ptr)` into a HIP graph on a side stream."""
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    exec_ = ctypes.c_void_p()
    with torch.cuda.stream(side):
        sp = side.cuda_stream
        nat.call('esr_graph_begin', sp)
        try:
            launch(sp)
        finally:
            rc = nat.lib().esr_graph_end(sp, ctypes.byref(exec_))
        nat.check(rc, 'esr_graph_end')
    torch.cuda.current_stream().wait_stream(side)
    return exec_


class _Graph:
    def __init__(self, launch):
        self.exec = _capture(launch)

    def replay(self):
        nat.call('esr_graph_launch', self.exec, nat.stream_ptr())

    def __del__(self):
        try:
            if self.exec:
                nat.lib().esr_graph_destroy(self.exec)
        except Exception:
            pass


class FusedCG:
    CHUNK = 32
    use_sell = True      # the product on the SELL-64 form (identical results)

    def __init__(self, cg, prog, op=None, diag=None):
        self.cg = cg
        self.op = LinsysCSR(op if op is not None else cg.A)
        self.diag = diag.data if diag is not None else None
        self.bound = BoundTensors(self.op.sources + [diag, cg.b, cg.x, cg.r, cg.z, cg.p])
        dev = cg.x.device
        n = self.op.n
        sb, wb = ctypes.c_int64(), ctypes.c_int64()
        nat.call('esr_pcg_workspace', n, ctypes.byref(sb), ctypes.byref(wb))
        self.state = torch.empty(sb.value, dtype=torch.uint8, device=dev)
        self.ws = torch.empty(wb.value, dtype=torch.uint8, device=dev)
        self.Ap = torch.empty(n, dtype=self.op.dtype, device=dev)
        self.dt = nat.dtype_code(self.op.dtype)
        # the SELL form is built on the first solve (set-up needs the GPU)
        self.want_sell = self.use_sell
        self.sell = None
        self.graph = None
        self.graph_chunk = 0

    def _arrays(self):
        c = self.cg
        return [self.op.rowptr.data_ptr(), self.op.col.data_ptr(),
                self.op.Ae.data_ptr(), self.op.Av.data_ptr(),
                None if self.diag is None else self.diag.data_ptr()]

    def steps(self, count, stream):
        c = self.cg
        vecs = [c.x.data.data_ptr(), c.r.data.data_ptr(), c.z.data.data_ptr(),
                c.p.data.data_ptr(), self.Ap.data_ptr(), self.state.data_ptr(),
                self.ws.data_ptr(), count, stream]
        if self.sell is not None:
            nat.call('esr_pcg_steps_sell', self.dt, self.op.n, *self.sell.arrays(),
                     self.op.Av.data_ptr(),
                     None if self.diag is None else self.diag.data_ptr(), *vecs)
        else:
            nat.call('esr_pcg_steps', self.dt, self.op.n, *self._arrays(), *vecs)

    def read(self):
        sc = (ctypes.c_double * 6)()
        it = (ctypes.c_int64 * 3)()
        nat.call('esr_pcg_read', self.state.data_ptr(), sc, it, nat.stream_ptr())
        return dict(rnorm=sc[0], bnorm=sc[1], rz=sc[2], tol=sc[3], iters=it[0],
                    done=it[1], k=it[2])

    def arm(self, rtol, atol=None, maxiter=None, chunk=None):
        """Init (cg.py:25-29) and the exit tolerance on the device; the
        chunk graph is captured on first use."""
        c = self.cg
        self.bound.check()
        if self.want_sell:
            self.want_sell = False
            self.sell = make_sell(self.op)
            self.graph = None
        elif self.sell is not None:
            self.sell.refresh(self.op.Ae)
        nat.call('esr_pcg_init', self.dt, self.op.n, *self._arrays(),
                 c.b.data.data_ptr(), c.x.data.data_ptr(), c.r.data.data_ptr(),
                 c.z.data.data_ptr(), c.p.data.data_ptr(), self.state.data_ptr(),
                 self.ws.data_ptr(), float(rtol), float(atol) if atol else 0.0,
                 -1 if maxiter is None else int(maxiter), nat.stream_ptr())
        chunk = chunk or self.CHUNK
        if self.graph is None or self.graph_chunk != chunk:
            self.graph = _Graph(lambda sp: self.steps(chunk, sp))
            self.graph_chunk = chunk

    def solve(self, rtol=1e-5, atol=None, maxiter=None, debug_iter=None,
              chunk=None):
        c = self.cg
        name = c.__class__.__name__
        self.arm(rtol, atol, maxiter, chunk or (debug_iter if debug_iter else None))
        while True:
            self.graph.replay()
            s = self.read()
            if debug_iter is not None:
                logger.info(f"{name} residual {s['rnorm']} at the "
                            f"{s['k']}-th iteration")
            if s['done']:
                break
        ops.fill_(c.rnorm.data, s['rnorm'])
        ops.fill_(c.bnorm.data, s['bnorm'])
        ops.fill_(c.rzsum.data, s['rz'])
        logger.info(f"{name} solver completed with residual {s['rnorm']}"
                    f" at the {s['iters']}-th iteration")
        return {'residual': s['rnorm'], 'iters': s['iters']}


def make_fused_cg(cg, prog):
    """The fused iteration when Step applies reducer(selector(p) * Ae) +
    p * Av and Update computes z = r or z = r / d (recognised in the traced
    graphs, whatever objects A and M are), else None."""
    from .recognize import cg_preconditioner, graph_of, match_linsys
    g_step, g_upd = graph_of(prog, cg.step), graph_of(prog, cg.update)
    op = match_linsys(cg.step, g_step, cg.p) if g_step is not None else None
    why = LinsysCSR.applicable(op, cg.x, prog)
    pre = cg_preconditioner(cg.update, g_upd, cg.r, cg.z) if g_upd is not None else None
    if why is None and pre is None:
        why = 'preconditioner is neither z = r nor z = r / d'
    if why is None and pre[0] == 'jacobi' and not _diag_ok(pre[1], cg.x, prog):
        why = 'the Jacobi diagonal is not a vector of x\'s space and dtype'
    if why is None and (cg.b.dtype != cg.x.dtype or cg.b.data.shape != cg.x.data.shape):
        why = 'b and x differ in dtype or shape'
    if why is not None:
        logger.info(f'CG: generic compiled path ({why})')
        return None
    return FusedCG(cg, prog, op, pre[1])


class FusedGMRES:
    def __init__(self, gm, prog, op=None, diag=None):
        self.gm = gm
        self.op = LinsysCSR(op if op is not None else gm.A)
        self.diag = diag.data if diag is not None else None
        self.bound = BoundTensors(self.op.sources + [diag, gm.b, gm.x, gm.r])
        dev = gm.x.device
        n, m = self.op.n, gm.restart
        sb, wb = ctypes.c_int64(), ctypes.c_int64()
        nat.call('esr_gmres_workspace', n, m, ctypes.byref(sb), ctypes.byref(wb))
        self.state = torch.empty(sb.value, dtype=torch.uint8, device=dev)
        self.ws = torch.empty(wb.value, dtype=torch.uint8, device=dev)
        self.V = torch.empty((m, n), dtype=self.op.dtype, device=dev)
        self.w = [torch.empty(n, dtype=self.op.dtype, device=dev) for _ in range(2)]
        self.y = torch.zeros(m, dtype=self.op.dtype, device=dev)
        self.dt = nat.dtype_code(self.op.dtype)
        self.graph = None
        self.loop_graph = None

    def _csr(self):
        return [self.op.rowptr.data_ptr(), self.op.col.data_ptr(),
                self.op.Ae.data_ptr(), self.op.Av.data_ptr()]

    def _cycle(self, sp):
        g = self.gm
        nat.call('esr_gmres_cycle', self.dt, self.op.n, g.restart, *self._csr(),
                 None if self.diag is None else self.diag.data_ptr(),
                 g.r.data.data_ptr(), self.V.data_ptr(), self.w[0].data_ptr(),
                 self.w[1].data_ptr(), self.state.data_ptr(), self.ws.data_ptr(), sp)

    def _read(self, with_H):
        m = self.gm.restart
        sc = (ctypes.c_double * 2)()
        it = (ctypes.c_int64 * 2)()
        H = np.zeros((m + 1) * m) if with_H else None
        nat.call('esr_gmres_read', self.state.data_ptr(), m, sc, it,
                 H.ctypes.data_as(_P_dbl) if with_H else None, nat.stream_ptr())
        return sc[0], sc[1], it[0], it[1], (H.reshape(m + 1, m) if with_H else None)

    # restart cycles per captured graph of the device loop
    LOOP_CYCLES = 4

    def _loop(self, sp):
        g = self.gm
        nat.call('esr_gmres_loop_cycles', self.dt, self.op.n, g.restart, *self._csr(),
                 None if self.diag is None else self.diag.data_ptr(), g.b.data.data_ptr(),
                 g.x.data.data_ptr(), g.r.data.data_ptr(), self.V.data_ptr(),
                 self.w[0].data_ptr(), self.w[1].data_ptr(), self.state.data_ptr(),
                 self.ws.data_ptr(), self.LOOP_CYCLES, sp)

    def _solve_device(self, rtol, atol, maxiter):
        """The restart loop on the device: graphs of LOOP_CYCLES cycles replayed
        until the device reports the exit (no host least squares)."""
        g = self.gm
        name = g.__class__.__name__
        self.bound.check()
        if self.loop_graph is None:
            self.loop_graph = _Graph(self._loop)
        sp = nat.stream_ptr()
        nat.call('esr_gmres_begin', self.dt, self.op.n, g.b.data.data_ptr(),
                 self.state.data_ptr(), self.ws.data_ptr(), sp)
        nat.call('esr_gmres_loop_arm', self.state.data_ptr(), float(rtol),
                 float(atol) if atol else 0.0, -1 if maxiter is None else int(maxiter), sp)
        sc = (ctypes.c_double * 2)()
        it = (ctypes.c_int64 * 2)()
        while True:
            self.loop_graph.replay()
            nat.call('esr_gmres_loop_read', self.state.data_ptr(), sc, it, nat.stream_ptr())
            if it[1]:
                break
        rnorm, bnorm, iters = sc[0], sc[1], int(it[0])
        _, _, _, _, H = self._read(True)
        g.easier_host_H[:] = H
        ops.fill_(g.rnorm.data, rnorm)
        ops.fill_(g.bnorm.data, bnorm)
        logger.info(f"{name} solver completed with residual {rnorm}"
                    f" at the {iters}-th iteration")
        return {'residual': rnorm, 'iters': iters}

    def solve(self, rtol=1e-5, atol=None, maxiter=None, debug_iter=None):
        """With debug_iter the restart loop runs on the host (residual logged
        every cycle, least squares by the reference's thin SVD); otherwise on
        the device."""
        if debug_iter is None:
            return self._solve_device(rtol, atol, maxiter)
        from .gmres import lstsq_svd
        g = self.gm
        name = g.__class__.__name__
        n, m = self.op.n, g.restart
        self.bound.check()
        if self.graph is None:
            self.graph = _Graph(self._cycle)
        nat.call('esr_gmres_begin', self.dt, n, g.b.data.data_ptr(),
                 self.state.data_ptr(), self.ws.data_ptr(), nat.stream_ptr())
        _, bnorm, _, _, _ = self._read(False)
        rtol = rtol * bnorm
        tol = max(rtol, atol) if atol else rtol
        iters = 0
        B = g.easier_host_B
        while True:
            nat.call('esr_gmres_residual', self.dt, n, *self._csr(),
                     g.b.data.data_ptr(), g.x.data.data_ptr(), g.r.data.data_ptr(),
                     self.state.data_ptr(), self.ws.data_ptr(), nat.stream_ptr())
            rnorm, _, _, _, _ = self._read(False)
            if debug_iter is not None and iters % debug_iter == 0:
                logger.info(f"{name} residual {rnorm} at the {iters}-th iteration")
            if (rnorm == rnorm and rnorm <= tol) or \
                    (maxiter is not None and iters >= maxiter):
                break
            if rnorm != rnorm and maxiter is None:
                logger.warning(f'{name}: residual is NaN, stopping')
                break
            iters += 1
            B[0, 0] = rnorm
            self.graph.replay()
            _, _, jend, broken, H = self._read(True)
            g.easier_host_H[:] = H
            y = lstsq_svd(H[:jend + 2], B[:jend + 2])
            ops.copy_(self.y, torch.from_numpy(y[:, 0].copy()).to(self.y.dtype)
                      .to(self.y.device))
            nat.call('esr_gmres_update_x', self.dt, n, jend + 1, self.V.data_ptr(),
                     self.y.data_ptr(),
                     None if self.diag is None else self.diag.data_ptr(),
                     g.x.data.data_ptr(), nat.stream_ptr())
        ops.fill_(g.rnorm.data, rnorm)
        ops.fill_(g.bnorm.data, bnorm)
        logger.info(f"{name} solver completed with residual {rnorm}"
                    f" at the {iters}-th iteration")
        return {'residual': rnorm, 'iters': iters}


def make_fused_gmres(gm, prog):
    """The fused Arnoldi cycles when UpdateRnorm applies
    reducer(selector(x) * Ae) + x * Av and UpdateX adds M(dx) = dx or
    dx / d (recognised in the traced graphs), else None."""
    from .recognize import gmres_preconditioner, graph_of, match_linsys
    g_res = graph_of(prog, gm.update_rnorm)
    op = match_linsys(gm.update_rnorm, g_res, gm.x) if g_res is not None else None
    why = LinsysCSR.applicable(op, gm.x, prog)
    ux = gm.update_x[0]
    g_ux = graph_of(prog, ux)
    pre = gmres_preconditioner(ux, g_ux, gm.x) if g_ux is not None else None
    if why is None and pre is None:
        why = 'preconditioner is neither M(v) = v nor M(v) = v / d'
    if why is None and pre[0] == 'jacobi' and not _diag_ok(pre[1], gm.x, prog):
        why = 'the Jacobi diagonal is not a vector of x\'s space and dtype'
    if why is None and gm.restart > 32:
        why = 'restart > 32'
    if why is not None:
        logger.info(f'GMRES: generic compiled path ({why})')
        return None
    return FusedGMRES(gm, prog, op, pre[1])
