"""Training and evaluation of the learned preconditioner (the configuration:
variable-coefficient P1 Poisson, log kappa a Gaussian random field).

Training pairs (r, e) come from the plain solver: Jacobi-PCG runs on seeded
problems from x = 0 with a uniform random right-hand side; at a schedule of
iterations k the residual r_k and the error e_k = x* - x_k (x* the converged
solution, relative residual 1e-10) are recorded. The network is trained on
the GPU with its own forward/backward kernels and an Adam step kernel to
minimise mean ||M(r) - e||^2 / ||e||^2. Everything runs on device buffers in
the mesh's natural order (no esr.compile renumbering is involved here).
"""
import ctypes
import math
import time
from typing import List

import numpy as np
import torch

from .. import _native as nat
from ..core import ops
from ..numeric.preconditioner.learned import LearnedPreconditioner, MeshGraph
from ..numeric.solver import fused as _fused  # noqa: F401  (registers esr_pcg_*)
from . import poisson as wp

SNAPSHOTS = (0, 1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024)


class Problem:
    """A variable-coefficient P1 Poisson problem on the GPU in natural order."""

    def __init__(self, n, seed, ell=0.05, jitter=0.0, device=None):
        dev = device or torch.device('cuda', torch.cuda.current_device())
        mesh, sys_, kap = wp.varcoef_system(n, seed, jitter=jitter, ell=ell)
        self.logk = wp.gaussian_random_field(n, seed, ell=ell).reshape(-1)
        self.mesh, self.sys, self.n_grid = mesh, sys_, n
        self.n = sys_['Av'].shape[0]
        rowptr = np.concatenate([[0], np.cumsum(np.bincount(sys_['row'], minlength=self.n))])
        i32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(dev)
        f64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)
        self.rowptr, self.col = i32(rowptr), i32(sys_['col'])
        self.Ae, self.Av, self.b = f64(sys_['Ae']), f64(sys_['Av']), f64(sys_['b'])
        self.graph = MeshGraph(sys_['row'], sys_['col'], sys_['Av'], sys_['Ae'],
                               mesh['points'], self.logk, n, dev)
        self.device = dev

    def csr(self):
        return [self.rowptr.data_ptr(), self.col.data_ptr(), self.Ae.data_ptr(),
                self.Av.data_ptr()]


class _Pcg:
    """Jacobi-PCG on a Problem through the fused kernels, with snapshots."""

    def __init__(self, prob: Problem):
        self.p = prob
        n, dev = prob.n, prob.device
        self.vec = {k: torch.zeros(n, dtype=torch.float64, device=dev)
                    for k in ('x', 'r', 'z', 'p', 'Ap')}
        sb, wb = ctypes.c_int64(), ctypes.c_int64()
        nat.call('esr_pcg_workspace', n, ctypes.byref(sb), ctypes.byref(wb))
        self.state = torch.empty(sb.value, dtype=torch.uint8, device=dev)
        self.ws = torch.empty(wb.value, dtype=torch.uint8, device=dev)

    def init(self, b, rtol, maxiter=None):
        v = self.vec
        ops.fill_(v['x'], 0)
        nat.call('esr_pcg_init', nat.F64, self.p.n, *self.p.csr(), self.p.Av.data_ptr(),
                 b.data_ptr(), v['x'].data_ptr(), v['r'].data_ptr(), v['z'].data_ptr(),
                 v['p'].data_ptr(), self.state.data_ptr(), self.ws.data_ptr(), rtol, 0.0,
                 -1 if maxiter is None else maxiter, nat.stream_ptr())

    def steps(self, count):
        v = self.vec
        nat.call('esr_pcg_steps', nat.F64, self.p.n, *self.p.csr(), self.p.Av.data_ptr(),
                 v['x'].data_ptr(), v['r'].data_ptr(), v['z'].data_ptr(), v['p'].data_ptr(),
                 v['Ap'].data_ptr(), self.state.data_ptr(), self.ws.data_ptr(), count,
                 nat.stream_ptr())

    def read(self):
        sc = (ctypes.c_double * 6)()
        it = (ctypes.c_int64 * 3)()
        nat.call('esr_pcg_read', self.state.data_ptr(), sc, it, nat.stream_ptr())
        return dict(rnorm=sc[0], bnorm=sc[1], iters=it[0], done=it[1], k=it[2])

    def solve(self, b, rtol=1e-8, chunk=64, maxiter=None):
        self.init(b, rtol, maxiter)
        while True:
            self.steps(chunk)
            s = self.read()
            if s['done']:
                return s


def generate_pairs(prob: Problem, n_rhs, seed, snapshots=SNAPSHOTS, store_dtype=torch.float32):
    """(r_k, e_k, ||e_k||^2) pairs from Jacobi-PCG runs with uniform random
    right-hand sides (seeded)."""
    pcg = _Pcg(prob)
    out = []
    gen = torch.Generator().manual_seed(seed)
    for k in range(n_rhs):
        b = (torch.rand(prob.n, generator=gen, dtype=torch.float64) * 2 - 1)
        b[torch.from_numpy(prob.sys['boundary'])] = 0.0
        bd = b.to(prob.device)
        pcg.init(bd, 1e-10, maxiter=max(20000, 4 * prob.n_grid * prob.n_grid))
        snaps = []
        done_at = 0
        for s in snapshots:
            pcg.steps(s - done_at)
            done_at = s
            st = pcg.read()
            if st['done']:
                break
            snaps.append((pcg.vec['r'].clone(), pcg.vec['x'].to(store_dtype)))
        while not pcg.read()['done']:
            pcg.steps(256)
        xstar = pcg.vec['x']
        for r, x in snaps:
            e = ops.ewise('sub', xstar, x.to(torch.float64))
            en2 = float(ops.aggregate('sum', ops.ewise('mul', e, e).reshape(-1, 1)).cpu())
            if en2 > 0:
                out.append((r, e.to(store_dtype), en2))
    return out


def train(pre: LearnedPreconditioner, data, epochs=1, lr=1e-3, seed=0, log_every=200,
          clip=1.0):
    """data: list of (Problem, pairs). Returns the per-step losses."""
    rng = np.random.default_rng(seed)
    losses = []
    step = 0
    for ep in range(epochs):
        order = [(i, j) for i, (_, pairs) in enumerate(data) for j in range(len(pairs))]
        rng.shuffle(order)
        order.sort(key=lambda t: t[0])           # one bind per problem per epoch
        bound = None
        for i, j in order:
            prob, pairs = data[i]
            if bound is not i:
                pre.bind(prob.graph, train=True)
                z = torch.empty(prob.n, dtype=torch.float64, device=prob.device)
                bound = i
            r, e, en2 = pairs[j]
            pre.apply(r, z)
            diff = ops.ewise('sub', z, e)
            dz = ops.ewise('mul', diff, 2.0 / en2)
            pre.backward(r, dz)
            pre.adam_step(lr=lr, clip=clip)
            step += 1
            if step % log_every == 0 or step == 1:
                l = float(ops.aggregate('sum', ops.ewise('mul', diff, diff).reshape(-1, 1))
                          .cpu()) / en2
                losses.append((step, l))
    return losses


def evaluate(pre: LearnedPreconditioner, prob: Problem, rtol=1e-8, seed=123,
             maxiter=None):
    """Wall-clock of flexible CG with the learned preconditioner against
    Jacobi-PCG (fused kernels) on the same right-hand side; both converge on
    the true float64 residual."""
    from ..numeric.solver.fcg import FcgKernels
    gen = torch.Generator().manual_seed(seed)
    b = torch.rand(prob.n, generator=gen, dtype=torch.float64) * 2 - 1
    b[torch.from_numpy(prob.sys['boundary'])] = 0.0
    bd = b.to(prob.device)
    pcg = _Pcg(prob)
    pcg.solve(bd, rtol)                      # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s = pcg.solve(bd, rtol)
    torch.cuda.synchronize()
    t_jac = time.perf_counter() - t0
    x_jac = pcg.vec['x'].clone()
    pre.bind(prob.graph)
    fk = FcgKernels(prob.rowptr, prob.col, prob.Ae, prob.Av, prob.n, prob.device)
    x = torch.zeros(prob.n, dtype=torch.float64, device=prob.device)
    fk.solve(bd, x, pre, rtol, maxiter=maxiter)   # warm-up
    ops.fill_(x, 0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    info = fk.solve(bd, x, pre, rtol, maxiter=maxiter)
    torch.cuda.synchronize()
    t_learn = time.perf_counter() - t0
    err = float(ops.aggregate('norm', ops.ewise('sub', x, x_jac).reshape(-1, 1)).cpu()) / \
        float(ops.aggregate('norm', x_jac.reshape(-1, 1)).cpu())
    return dict(n=prob.n, jacobi_iters=s['iters'], jacobi_s=t_jac,
                jacobi_residual=s['rnorm'], learned_iters=info['iters'], learned_s=t_learn,
                learned_residual=info['residual'], fallback_at=info['fallback_at'],
                speedup=t_jac / t_learn, rel_diff_x=err, bnorm=s['bnorm'])


def jacobi_loss(data):
    """The same loss for the untrained (Jacobi) preconditioner z = r/d."""
    tot, cnt = 0.0, 0
    for prob, pairs in data:
        for r, e, en2 in pairs:
            z = ops.ewise('div', r, prob.Av)
            d = ops.ewise('sub', z, e)
            tot += float(ops.aggregate('sum', ops.ewise('mul', d, d).reshape(-1, 1)).cpu()) / en2
            cnt += 1
    return tot / max(cnt, 1)
