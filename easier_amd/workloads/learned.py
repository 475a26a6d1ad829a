"""Training and evaluation of the learned preconditioner (the configuration:
variable-coefficient P1 Poisson, log kappa a Gaussian random field).

Training pairs (r, e) come from the plain solver: Jacobi-PCG runs on seeded
problems from x = 0 with a uniform random right-hand side; at a schedule of
iterations k the residual r_k and the error e_k = x* - x_k (x* the converged
solution, relative residual 1e-10) are recorded. The network is trained on
the GPU with its own forward/backward kernels and an Adam step kernel. The
default loss is the A-norm error left after the best step along z = M(r),
1 - (z.r)^2 / ((z.Az) ||e||_A^2), which is what one flexible-CG step gains;
the plain error loss ||M(r) - e||^2 / ||e||^2 is the alternative. Training
runs on device buffers in the mesh's natural order; evaluation goes through
esr.compile like any user program.
"""
import ctypes
import math
import time
from typing import List, Optional

import numpy as np
import torch

from .. import _native as nat
from ..core import ops
from ..numeric.preconditioner.learned import GnnBuffers, LearnedPreconditioner, MeshGraph
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


def _dot(a, b):
    return float(ops.aggregate('sum', ops.ewise('mul', a, b).reshape(-1, 1)).cpu())


def spmv(prob: Problem, x: torch.Tensor) -> torch.Tensor:
    """A x with the problem's operator (Linsys: off-diagonal CSR + diagonal)."""
    g = ops.ewise('mul', ops.gather(x, prob.col), prob.Ae)
    off = ops.csr_reduce(g, prob.rowptr, None, None, prob.n, 'sum')
    return ops.ewise('add', off, ops.ewise('mul', x, prob.Av))


def generate_pairs(prob: Problem, n_rhs, seed, snapshots=SNAPSHOTS, store_dtype=torch.float64):
    """(r_k, e_k, ||e_k||^2, r_k.e_k) from Jacobi-PCG runs with uniform
    random right-hand sides (seeded); r_k.e_k = ||e_k||_A^2 since A e_k = r_k.
    The iterates are kept in float64: e_k = x* - x_k of a late snapshot is
    far smaller than x_k, and float32 rounding of x_k would put
    high-frequency noise into e_k that A e_k amplifies."""
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
            snaps.append((pcg.vec['r'].clone(), pcg.vec['x'].to(store_dtype, copy=True)))
        while not pcg.read()['done']:
            pcg.steps(256)
        xstar = pcg.vec['x']
        for r, x in snaps:
            e = ops.ewise('sub', xstar, x.to(torch.float64))
            en2 = _dot(e, e)
            if en2 > 0:
                out.append((r, e.to(store_dtype), en2, _dot(r, e)))
    return out


def generate_pairs_random(prob: Problem, count, seed, store_dtype=torch.float32):
    """(r, e, ||e||^2, r.e) with e uniform[-1, 1] on the interior vertices
    (zero on the boundary, seeded) and r = A e: errors with every frequency,
    the part a smoother has to remove."""
    gen = torch.Generator().manual_seed(seed)
    inner = torch.from_numpy(~prob.sys['boundary'])
    out = []
    for _ in range(count):
        e = (torch.rand(prob.n, generator=gen, dtype=torch.float64) * 2 - 1) * inner
        ed = e.to(prob.device)
        r = spmv(prob, ed)
        out.append((r, ed.to(store_dtype), _dot(ed, ed), _dot(r, ed)))
    return out


def generate_pairs_fcg(prob: Problem, pre: LearnedPreconditioner, n_rhs, seed,
                       every=32, max_snaps=24, store_dtype=torch.float32):
    """Pairs from the iterates of flexible CG preconditioned by the network
    itself (the residuals the network meets in use): the same right-hand
    sides as generate_pairs' seed, a snapshot every `every` steps."""
    from ..numeric.solver.fcg import RESIDUAL as RESIDUAL_OP, FcgKernels
    fk = FcgKernels(prob.rowptr, prob.col, prob.Ae, prob.Av, prob.n, prob.device,
                    chunk=every)
    pre.bind(prob.graph)
    pcg = _Pcg(prob)
    out = []
    gen = torch.Generator().manual_seed(seed)
    x = torch.zeros(prob.n, dtype=torch.float64, device=prob.device)
    for k in range(n_rhs):
        b = (torch.rand(prob.n, generator=gen, dtype=torch.float64) * 2 - 1)
        b[torch.from_numpy(prob.sys['boundary'])] = 0.0
        bd = b.to(prob.device)
        ops.fill_(x, 0)
        fk.op(RESIDUAL_OP, b=bd, x=x)
        tol = 1e-10 * float(np.sqrt(fk.read()['bb']))
        nat.call('esr_fcg_set', fk.state.data_ptr(), tol * tol, -1, nat.stream_ptr())
        snaps = [(fk.r.clone(), x.to(store_dtype, copy=True))]
        fk._step(pre, x, True)
        while len(snaps) < max_snaps:
            fk._chunk_graph(pre, x).replay()
            s = fk.read()
            if s['done']:
                break
            snaps.append((fk.r.clone(), x.to(store_dtype, copy=True)))
        pcg.init(bd, 1e-10, maxiter=max(20000, 4 * prob.n_grid * prob.n_grid))
        while not pcg.read()['done']:
            pcg.steps(256)
        xstar = pcg.vec['x']
        for r, xs in snaps:
            e = ops.ewise('sub', xstar, xs.to(torch.float64))
            en2 = _dot(e, e)
            if en2 > 0:
                out.append((r, e.to(store_dtype), en2, _dot(r, e)))
    return out


class TwoGrid:
    """The two-grid cycle a smoother is trained for: on a Problem's
    multigrid hierarchy, the error left after smoothing with z and the exact
    coarse-grid correction is Q (e - z), Q = I - P Ac^-1 P^T A (the
    A-orthogonal projector onto the complement of the coarse space). Ac^-1 is
    applied as `coarse_cycles` stationary V-cycles (damped-Jacobi smoothing)
    from the coarse level down."""

    def __init__(self, prob: Problem, coarse_cycles=4, omega=0.8):
        from ..numeric.preconditioner.multigrid import MultigridPreconditioner
        coarsest = min(16, max(1, prob.n_grid // 2))
        self.mg = MultigridPreconditioner(None, omega, coarsest=coarsest,
                                         coarsening='geometric').setup(
            prob.rowptr, prob.col, prob.Ae, prob.Av, prob.n_grid)
        if len(self.mg.levels) < 2:
            raise ValueError(f'a {prob.n_grid} x {prob.n_grid} mesh has no coarse level')
        self.cycles = coarse_cycles
        dev = prob.device
        n, nc = self.mg.levels[0].nv, self.mg.levels[1].nv
        mk = lambda m: torch.zeros(m, dtype=torch.float64, device=dev)
        self.t, self.y, self.aqw, self.zero = mk(n), mk(n), mk(n), mk(n)
        self.rc, self.ec, self.res, self.de = mk(nc), mk(nc), mk(nc), mk(nc)

    @staticmethod
    def _csr(L):
        return (L.rowptr.data_ptr(), L.col.data_ptr(), L.Ae.data_ptr(), L.Av.data_ptr())

    def coarse_solve(self):
        """ec ~ Ac^-1 rc."""
        L1 = self.mg.levels[1]
        sp = nat.stream_ptr()
        self.mg._cycle(1, self.rc, self.ec)
        for _ in range(self.cycles - 1):
            nat.call('esr_mg_residual', nat.F64, L1.nv, L1.n_grid, None, *self._csr(L1),
                     self.rc.data_ptr(), self.ec.data_ptr(), None, None, self.res.data_ptr(), sp)
            self.mg._cycle(1, self.res, self.de)
            ops.ewise('add', self.ec, self.de, out=self.ec)

    def project(self, t):
        """A Q v for t = A v (into self.aqw; self.y = P Ac^-1 P^T t)."""
        L0 = self.mg.levels[0]
        sp = nat.stream_ptr()
        nat.call('esr_mg_restrict', nat.F64, L0.n_grid, None, t.data_ptr(),
                 self.rc.data_ptr(), sp)
        self.coarse_solve()
        nat.call('esr_mg_residual', nat.F64, L0.nv, L0.n_grid, L0.par.data_ptr(),
                 *self._csr(L0), t.data_ptr(), self.zero.data_ptr(), self.ec.data_ptr(),
                 self.y.data_ptr(), self.aqw.data_ptr(), sp)
        return self.aqw

    def aq(self, r, z):
        """A Q (e - z) for A e = r (into self.aqw)."""
        L0 = self.mg.levels[0]
        nat.call('esr_mg_residual', nat.F64, L0.nv, L0.n_grid, None, *self._csr(L0),
                 r.data_ptr(), z.data_ptr(), None, None, self.t.data_ptr(), nat.stream_ptr())
        return self.project(self.t)


def pair_loss(prob: Problem, r, e, en2, re, z, kind='energy'):
    """(loss, dloss/dz) of the preconditioned vector z = M(r).

    kind 'twogrid': ||Q (e - z)||_A^2 / ||e||_A^2, the A-norm error left by a
        two-grid cycle that pre-smooths with z (TwoGrid; prob.twogrid); its
        gradient is -2 A Q (e - z) / ||e||_A^2 (Q is A-orthogonal). Not
        scale-invariant: the smoother is used with unit step.
    kind 'twogrid_post': the same cycle followed by damped-Jacobi
        post-smoothing S = I - w D^-1 A (w = 0.8), as the multigrid
        preconditioner runs it: ||S Q (e - z)||_A^2 / ||e||_A^2, gradient
        -2 A Q S^2 Q (e - z) / ||e||_A^2 (S and Q are A-self-adjoint).

    kind 'error':  ||z - e||^2 / ||e||^2.
    kind 'energy': 1 - (z.r)|z.r| / ((z.Az) (r.e)) -- for z.r > 0 the
        fraction of the A-norm error left after the best step along z
        (r.e = ||e||_A^2); above 1 when z is not a descent direction. Its
        gradient is -2 |z.r| / ((z.Az)(r.e)) (r - (z.r)/(z.Az) A z).
        Scale-invariant in z, like the step flexible CG takes."""
    if kind == 'error':
        diff = ops.ewise('sub', z, e)
        return _dot(diff, diff) / en2, ops.ewise('mul', diff, 2.0 / en2)
    if kind in ('twogrid', 'twogrid_post'):
        if getattr(prob, 'twogrid', None) is None:
            prob.twogrid = TwoGrid(prob)
        tg = prob.twogrid
        aqw = tg.aq(r, z)
        w = ops.ewise('sub', e.to(torch.float64), z)
        if kind == 'twogrid':
            return _dot(w, aqw) / re, ops.ewise('mul', aqw, -2.0 / re)
        omega = 0.8
        u = ops.ewise('sub', w, tg.y)                                  # Q w
        s1 = ops.ewise('sub', u, ops.ewise('mul', ops.ewise('div', aqw, prob.Av), omega))
        as1 = spmv(prob, s1)
        loss = _dot(s1, as1) / re
        s2 = ops.ewise('sub', s1, ops.ewise('mul', ops.ewise('div', as1, prob.Av), omega))
        aq2 = tg.project(spmv(prob, s2))
        return loss, ops.ewise('mul', aq2, -2.0 / re)
    q = spmv(prob, z)
    zr, zq = _dot(z, r), _dot(z, q)
    loss = 1.0 - zr * abs(zr) / (zq * re)
    t = ops.ewise('sub', r, ops.ewise('mul', q, zr / zq))
    return loss, ops.ewise('mul', t, -2.0 * abs(zr) / (zq * re))


def train(pre: LearnedPreconditioner, data, epochs=1, lr=1e-3, seed=0, log_every=200,
          clip=1.0, loss='energy', log=None):
    """data: list of (Problem, pairs). Returns [(step, loss)] every
    `log_every` steps (and the mean loss of each epoch as (step, mean));
    log(step, running mean) is called at the same steps."""
    rng = np.random.default_rng(seed)
    losses = []
    step = 0
    for ep in range(epochs):
        order = [(i, j) for i, (_, pairs) in enumerate(data) for j in range(len(pairs))]
        rng.shuffle(order)
        order.sort(key=lambda t: t[0])           # one bind per problem per epoch
        bound = None
        tot = 0.0
        for k, (i, j) in enumerate(order, 1):
            prob, pairs = data[i]
            if bound != i:
                pre.bind(prob.graph, train=True)
                z = torch.empty(prob.n, dtype=torch.float64, device=prob.device)
                bound = i
            r, e, en2, re = pairs[j]
            pre.apply(r, z)
            l, dz = pair_loss(prob, r, e, en2, re, z, loss)
            pre.backward(r, dz)
            pre.adam_step(lr=lr, clip=clip)
            step += 1
            tot += l
            if step % log_every == 0 or step == 1:
                losses.append((step, l))
                if log is not None:
                    log(step, tot / k)
        losses.append((step, tot / max(len(order), 1)))
    return losses


def mean_loss(pre: Optional[LearnedPreconditioner], data, loss='energy', omega=1.0):
    """Mean pair_loss over the pairs; pre None = the Jacobi preconditioner
    z = omega r / d (omega = 1: the untrained network)."""
    tot, cnt = 0.0, 0
    for prob, pairs in data:
        if pre is not None:
            pre.bind(prob.graph)
        for r, e, en2, re in pairs:
            z = ops.ewise('mul', ops.ewise('div', r, prob.Av), omega) if pre is None \
                else pre(r)
            tot += pair_loss(prob, r, e, en2, re, z, loss)[0]
            cnt += 1
    return tot / max(cnt, 1)


METHODS = ('jacobi', 'learned', 'mg', 'mg-learned', 'amg', 'amg-learned')


def make_preconditioner(method, pre: Optional[LearnedPreconditioner], **mg_args):
    """The preconditioner of flexible CG for an evaluation method: 'learned'
    = the network alone, 'mg' = multigrid V-cycle with damped-Jacobi
    smoothing (geometric coarsening where it applies), 'mg-learned' = the
    same cycle with the network as smoother; 'amg' / 'amg-learned' = the
    same with the hierarchy built by smoothed aggregation (the network on
    the finest level)."""
    from ..numeric.preconditioner.multigrid import MultigridPreconditioner
    if method == 'learned':
        return pre
    if method in ('mg', 'amg'):
        kind = 'aggregation' if method == 'amg' else 'auto'
        return MultigridPreconditioner(None, coarsening=kind, **mg_args)
    if method in ('mg-learned', 'amg-learned'):
        kind = 'aggregation' if method == 'amg-learned' else 'auto'
        return MultigridPreconditioner(pre, coarsening=kind, **mg_args)
    raise ValueError(f'unknown method {method!r}')


def evaluate_compiled(pre: Optional[LearnedPreconditioner], n, seed, rtol=1e-8, maxiter=None,
                      reps=2, methods=('jacobi', 'learned'), mg_args=None):
    """The drop-in comparison on the variable-coefficient problem of `seed`
    on the n x n mesh: each method compiled with esr.compile(..., 'hip') --
    'jacobi' is the fused Jacobi-PCG (CG solver), the others flexible CG with
    make_preconditioner(method) -- all from x = 0 to `rtol` relative
    residual. Solve times are the best of `reps` solves after one warm-up;
    `setup_seconds` is the compile (renumbering, device transfer and, for
    multigrid, the hierarchy). The true float64 residual ||b - A x|| of each
    final iterate is recomputed on the host."""
    import scipy.sparse as sp
    from .. import compile as esr_compile
    from ..numeric.solver import CG, FlexibleCG
    mesh, sys_, _ = wp.varcoef_system(n, seed)
    logk = wp.gaussian_random_field(n, seed).reshape(-1)
    nv = sys_['Av'].shape[0]
    A = sp.csr_matrix((sys_['Ae'], (sys_['row'], sys_['col'])), shape=(nv, nv)) + \
        sp.diags(sys_['Av'])
    bnorm = float(np.linalg.norm(sys_['b']))
    out = dict(n_grid=n, vertices=nv, seed=seed, rtol=rtol)

    def timed(solve, x):
        best, info = None, None
        for rep in range(reps + 1):
            ops.fill_(x.data, 0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            info = solve()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            if rep > 0:
                best = dt if best is None else min(best, dt)
        return best, info

    xs = {}
    for method in methods:
        eqn = wp.PoissonP1(sys_)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if method == 'jacobi':
            sol = CG(eqn.A, eqn.b, eqn.x, eqn.M)
        else:
            M = make_preconditioner(method, pre, **(mg_args or {}))
            sol = FlexibleCG(eqn.A, eqn.b, eqn.x, M,
                             dict(points=mesh['points'], logk=logk, n_grid=n))
        esr_compile([eqn, sol], 'hip')
        torch.cuda.synchronize()
        setup = time.perf_counter() - t0
        t, info = timed(lambda: sol.solve(rtol=rtol, maxiter=maxiter), eqn.x)
        x = eqn.x.collect().numpy()
        xs[method] = x
        out[method] = dict(seconds=t, iters=info['iters'], setup_seconds=setup,
                           fallback_at=info.get('fallback_at'),
                           true_rel_residual=float(np.linalg.norm(sys_['b'] - A @ x)) / bnorm)
        _log(f'{method}: {out[method]}')
        del eqn, sol
    if 'jacobi' in out:
        for method in methods:
            if method != 'jacobi':
                r = out[method]
                r['speedup_vs_jacobi'] = out['jacobi']['seconds'] / r['seconds']
                r['iteration_ratio'] = out['jacobi']['iters'] / max(r['iters'], 1)
                r['rel_diff_x'] = float(np.linalg.norm(xs[method] - xs['jacobi']) /
                                        np.linalg.norm(xs['jacobi']))
    return out


def _log(msg):
    import sys
    print(f'[learned {time.strftime("%H:%M:%S")}] {msg}', file=sys.stderr, flush=True)


def time_forward(pre: LearnedPreconditioner, n, seed=1, reps=50):
    """Microseconds per network application z = M(r) (||r||, encoder and the
    layers) on the n x n problem of `seed`, HIP events over `reps`
    applications after 3 warm-up ones, r uniform in [0, 1)."""
    prob = Problem(n, seed)
    pre.bind(prob.graph)
    bufs = GnnBuffers(prob.n, prob.device)
    g = torch.Generator(device='cpu').manual_seed(0)
    r = torch.rand(prob.n, generator=g, dtype=torch.float64).to(prob.device)
    z = torch.empty_like(r)
    for _ in range(3):
        pre.forward_into(prob.graph, bufs, r, z)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        pre.forward_into(prob.graph, bufs, r, z)
    b.record()
    torch.cuda.synchronize()
    return dict(n_grid=n, vertices=prob.n, edges=int(prob.col.shape[0]), layers=pre.L,
                seed=seed, reps=reps, us_per_application=a.elapsed_time(b) * 1e3 / reps)


def main(argv=None):
    """python -m easier_amd.workloads.learned train|evaluate|forward ..."""
    import argparse
    from ..numeric.preconditioner import learned as lp
    ap = argparse.ArgumentParser(prog='python -m easier_amd.workloads.learned')
    sub = ap.add_subparsers(dest='cmd', required=True)
    tr = sub.add_parser('train', help='generate (r, e) pairs with Jacobi-PCG on seeded '
                        'problems and train the network on the GPU')
    tr.add_argument('--n', type=int, nargs='+', default=[512],
                    help='mesh sizes; problems are dealt over them in turn')
    tr.add_argument('--problems', type=int, default=24)
    tr.add_argument('--rhs', type=int, default=36)
    tr.add_argument('--epochs', type=int, default=4)
    tr.add_argument('--lr', type=float, default=1e-3)
    tr.add_argument('--layers', type=int, default=3)
    tr.add_argument('--seed0', type=int, default=1000)
    tr.add_argument('--loss', choices=['energy', 'error', 'twogrid', 'twogrid_post'],
                    default='twogrid_post')
    tr.add_argument('--policy-rounds', type=int, default=0,
                    help='rounds of pairs from flexible CG with the network itself')
    tr.add_argument('--policy-rhs', type=int, default=8)
    tr.add_argument('--random-pairs', type=int, default=0,
                    help='pairs per problem with a uniform random error e and r = A e')
    tr.add_argument('--out', required=True)
    ev = sub.add_parser('evaluate', help='learned FCG vs Jacobi-PCG on unseen seeds')
    ev.add_argument('--weights', required=True,
                    help="an .npz from 'train', or 'untrained' (the network equal to Jacobi)")
    ev.add_argument('--n', type=int, default=2048)
    ev.add_argument('--seeds', type=int, nargs='+', default=[1])
    ev.add_argument('--rtol', type=float, default=1e-8)
    ev.add_argument('--maxiter', type=int, default=50000)
    ev.add_argument('--methods', nargs='+', choices=METHODS, default=list(METHODS))
    ev.add_argument('--learned-min-grid', type=int, default=64,
                    help='multigrid levels with at least this many cells per side use '
                         'the network as smoother')
    ev.add_argument('--learned-smoothing', choices=['pre', 'post', 'both'], default='pre',
                    help='where the network smooths on those levels (damped Jacobi elsewhere)')
    fw = sub.add_parser('forward', help='time network applications z = M(r) on the '
                        'n x n problem of a seed (vertices in natural order)')
    fw.add_argument('--weights', required=True)
    fw.add_argument('--n', type=int, default=2048)
    fw.add_argument('--seed', type=int, default=1)
    fw.add_argument('--reps', type=int, default=50)
    args = ap.parse_args(argv)
    import json
    if args.cmd == 'forward':
        print(json.dumps(time_forward(lp.load(args.weights), args.n, args.seed, args.reps)),
              flush=True)
        return
    if args.cmd == 'train':
        data, t0 = [], time.perf_counter()
        sizes = [args.n[k % len(args.n)] for k in range(args.problems)]
        for k in range(args.problems):
            prob = Problem(sizes[k], args.seed0 + k)
            pairs = generate_pairs(prob, args.rhs, seed=args.seed0 + k)
            if args.random_pairs:
                pairs += generate_pairs_random(prob, args.random_pairs,
                                               seed=2 * 10 ** 6 + args.seed0 + k)
            data.append((prob, pairs))
            _log(f'problem {k + 1}/{args.problems} ({sizes[k]}^2, seed {args.seed0 + k}): '
                 f'{len(pairs)} pairs, {time.perf_counter() - t0:.1f} s')
        npairs = sum(len(p) for _, p in data)
        base = mean_loss(None, data, args.loss)
        base08 = mean_loss(None, data, args.loss, omega=0.8)
        _log(f'{npairs} pairs; Jacobi {args.loss} loss {base:.4f} (damped 0.8: {base08:.4f})')
        pre = LearnedPreconditioner(args.layers, lp.init_params(args.layers, args.seed0))

        def epochs(tag):
            for ep in range(args.epochs):
                lr = args.lr * (0.5 ** ep)
                losses = train(pre, data, epochs=1, lr=lr, seed=ep, log_every=500,
                               loss=args.loss,
                               log=lambda st, m: _log(f'{tag} step {st}: running mean {m:.4f}'))
                _log(f'{tag} epoch {ep + 1}/{args.epochs} lr {lr:.2e}: mean '
                     f'{losses[-1][1]:.4f} ({time.perf_counter() - t0:.1f} s)')
        epochs('pcg pairs')
        for rnd in range(args.policy_rounds):
            # residuals of flexible CG with the current network, added to the set
            for k, (prob, pairs) in enumerate(data):
                pairs.extend(generate_pairs_fcg(prob, pre, args.policy_rhs,
                                                seed=10 ** 6 + args.seed0 + 97 * rnd + k))
            npairs = sum(len(p) for _, p in data)
            _log(f'round {rnd + 1}: {npairs} pairs with flexible-CG iterates '
                 f'({time.perf_counter() - t0:.1f} s)')
            epochs(f'round {rnd + 1}')
        final = mean_loss(pre, data, args.loss)
        _log(f'trained {args.loss} loss {final:.4f} (Jacobi {base:.4f}, damped 0.8 '
             f'{base08:.4f})')
        lp.save(args.out, pre, n_train=np.asarray(sizes), pairs=npairs, jacobi_loss=base,
                damped_jacobi_loss=base08, trained_loss=final, loss=args.loss,
                seeds=np.arange(args.seed0, args.seed0 + args.problems))
        _log(f'saved {args.out}')
    else:
        pre = lp.load(args.weights) if args.weights != 'untrained' else \
            LearnedPreconditioner(3, lp.init_params(3, 0))
        for s in args.seeds:
            res = evaluate_compiled(pre, args.n, s, args.rtol, maxiter=args.maxiter,
                                    methods=args.methods,
                                    mg_args=dict(learned_min_grid=args.learned_min_grid,
                                                 learned_pre=args.learned_smoothing != 'post',
                                                 learned_post=args.learned_smoothing != 'pre'))
            print(json.dumps(res), flush=True)


if __name__ == '__main__':
    main()
