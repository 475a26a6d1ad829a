"""Learned preconditioner z = M(r): an edge message-passing network on the
mesh graph, executed by hand-written kernels (csrc/k_gnn.hip).

No reference counterpart (easier/numeric/preconditioner/__init__.py:4 is an
empty slot; CG/GMRES accept any callable M, cg.py:86). The network is
nonlinear in r, so it is used inside flexible CG (solver/fcg.py).

Parameters live in one flat float32 device tensor (layout below, shared with
the kernels); the matrices the matrix cores consume are packed to bfloat16
after every update. GnnApply wraps forward + backward as a
torch.autograd.Function (float32 gradients w.r.t. the flat parameters).

Flat layout (H = 32): W_in (3, H), b_in (H); per layer Wp (H, H), Wq (H, H),
We (3, H), bm (H), Wu (2H, H), bu (H); w_out (H), b_out (1). Matrices are
row-major with the input index first (y = x W).
"""
import ctypes
from typing import Dict, Optional

import numpy as np
import torch

from ... import _native as nat
from ...core import ops

H = 32
_c_i32, _c_i64, _c_vp, _c_dbl = ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_double
nat.register_signatures({
    'esr_gnn_sizes': (_c_i32, [_c_i32, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_i64)]),
    'esr_gnn_pack': (_c_i32, [_c_i32, _c_vp, _c_vp, _c_vp]),
    'esr_gnn_forward': (_c_i32, [_c_i32, _c_i64] + [_c_vp] * 12 + [_c_i32, _c_vp, _c_vp]),
    'esr_gnn_forward_slab': (_c_i32, [_c_i32, _c_i64] + [_c_vp] * 13 + [_c_i32, _c_vp, _c_vp]),
    'esr_gnn_backward': (_c_i32, [_c_i32, _c_i64, _c_i64] + [_c_vp] * 16 + [_c_i64, _c_vp]),
    'esr_gnn_adam': (_c_i32, [_c_i64, _c_vp, _c_vp, _c_vp, _c_vp, _c_dbl, _c_dbl, _c_dbl,
                              _c_dbl, _c_i64, _c_dbl, _c_vp]),
})


def param_shapes(L):
    shapes = [('W_in', (3, H)), ('b_in', (H,))]
    for l in range(L):
        shapes += [(f'Wp{l}', (H, H)), (f'Wq{l}', (H, H)), (f'We{l}', (3, H)),
                   (f'bm{l}', (H,)), (f'Wu{l}', (2 * H, H)), (f'bu{l}', (H,))]
    return shapes + [('w_out', (H,)), ('b_out', ())]


def init_params(L=3, seed=0) -> Dict[str, torch.Tensor]:
    """Initial float32 parameters for training: weights ~ N(0, 1/fan_in)
    (Wu halved, so each residual layer starts as a small update), zero
    biases and a zero decoder, so the untrained network is exactly Jacobi."""
    rng = np.random.default_rng(seed)

    def w(rows, cols, scale=1.0):
        return torch.from_numpy((rng.standard_normal((rows, cols)) * scale /
                                 np.sqrt(rows)).astype(np.float32))
    p = {'W_in': w(3, H), 'b_in': torch.zeros(H)}
    for l in range(L):
        p[f'Wp{l}'], p[f'Wq{l}'], p[f'We{l}'] = w(H, H), w(H, H), w(3, H)
        p[f'bm{l}'] = torch.zeros(H)
        p[f'Wu{l}'], p[f'bu{l}'] = w(2 * H, H, 0.5), torch.zeros(H)
    p['w_out'], p['b_out'] = torch.zeros(H), torch.zeros(())
    return p


def save(path, pre: 'LearnedPreconditioner', **meta):
    """Parameters as an .npz of plain arrays (loadable without pickle)."""
    np.savez(path, params=pre.params.detach().cpu().numpy(), L=np.int64(pre.L),
             H=np.int64(H), **{k: np.asarray(v) for k, v in meta.items()})


def load(path, device=None) -> 'LearnedPreconditioner':
    with np.load(path, allow_pickle=False) as f:
        if int(f['H']) != H:
            raise ValueError(f'{path}: hidden width {int(f["H"])}, this build has {H}')
        L = int(f['L'])
        flat = torch.from_numpy(f['params'].astype(np.float32))
    return LearnedPreconditioner(L, unflatten(flat, L), device)


def flatten(params: Dict[str, torch.Tensor], L) -> torch.Tensor:
    return torch.cat([params[k].reshape(-1).float() for k, _ in param_shapes(L)])


def unflatten(flat: torch.Tensor, L) -> Dict[str, torch.Tensor]:
    out, o = {}, 0
    for k, shp in param_shapes(L):
        sz = int(np.prod(shp)) if shp else 1
        out[k] = flat[o:o + sz].reshape(shp)
        o += sz
    return out


class EdgeSlab:
    """The graph's edges for the network's forward kernels, slot-major per
    group of 16 vertices (the vertices of one wavefront): edge k of vertex
    16 q + j at slot soff[q] + 16 k + j, group width the group's largest
    degree; ecnt the degrees (uint8), ecol the neighbour of each slot and
    feat its 3 features (padded slots: the slot's own vertex, zero
    features). The same edges in the same order as the CSR (rowptr, col,
    efeat) it is built from. Raises ValueError for a degree above 255."""

    def __init__(self, rowptr, col, efeat, n, device):
        rp = np.asarray(rowptr, np.int64)
        col = np.asarray(col, np.int64)
        efeat = np.asarray(efeat, np.float32).reshape(-1, 3)
        deg = np.diff(rp)
        if deg.size and deg.max() > 255:
            raise ValueError('a vertex has more than 255 edges')
        ng = (n + 15) // 16
        dp = np.zeros(ng * 16, np.int64)
        dp[:n] = deg
        width = dp.reshape(ng, 16).max(axis=1)
        soff = np.zeros(ng + 1, np.int64)
        soff[1:] = np.cumsum(16 * width)
        slots = int(soff[-1])
        # the vertex each slot belongs to (padding rows past n: vertex n - 1)
        base = np.repeat(soff[:-1], 16 * width)
        vert = np.repeat(np.arange(ng, dtype=np.int64) * 16, 16 * width) + \
            (np.arange(slots, dtype=np.int64) - base) % 16
        ecol = np.minimum(vert, n - 1)
        rows = np.repeat(np.arange(n, dtype=np.int64), deg)
        k = np.arange(rp[-1], dtype=np.int64) - rp[rows]
        slot = soff[rows // 16] + 16 * k + rows % 16
        ecol[slot] = col
        feat = np.zeros((max(slots, 1), 3), np.float32)
        feat[slot] = efeat
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
        self.soff = t(soff)
        self.ecnt = t(deg.astype(np.uint8))
        self.ecol = t(np.concatenate([ecol, np.zeros(1 if slots == 0 else 0, np.int64)])
                      .astype(np.int32))
        self.feat = t(feat)
        self.slots = slots

    @staticmethod
    def of(graph):
        """EdgeSlab of a graph with n, rowptr, col, efeat (device tensors)."""
        return EdgeSlab(graph.rowptr.cpu().numpy(), graph.col.cpu().numpy(),
                        graph.efeat.cpu().numpy(), graph.n, graph.rowptr.device)


def _forward_call(pre, g, r, rnorm, h, Q, a, save, z):
    """esr_gnn_forward on the graph's slab when it has one, else its CSR."""
    tail = [g.vfeat.data_ptr(), r.data_ptr(), g.d.data_ptr(), rnorm.data_ptr(),
            pre.params.data_ptr(), pre.packed.data_ptr(), h.data_ptr(), Q.data_ptr(),
            None if a is None else a.data_ptr(), save, z.data_ptr(), nat.stream_ptr()]
    slab = getattr(g, 'slab', None)
    if slab is not None:
        nat.call('esr_gnn_forward_slab', pre.L, g.n, slab.soff.data_ptr(),
                 slab.ecnt.data_ptr(), slab.ecol.data_ptr(), slab.feat.data_ptr(), *tail)
    else:
        nat.call('esr_gnn_forward', pre.L, g.n, g.rowptr.data_ptr(), g.col.data_ptr(),
                 g.efeat.data_ptr(), *tail)


class MeshGraph:
    """Static inputs of the network for one operator: the CSR (row-sorted
    directed vertex edges), the reverse edge of every edge, vertex features
    (log(d/mean d), log kappa) and edge features (a_ij/sqrt(d_i d_j), edge
    vector in grid units), and the diagonal d."""

    def __init__(self, row, col, Av, Ae, points, logk_vertex, n_grid, device=None,
                 use_slab=True):
        dev = device or torch.device('cuda', torch.cuda.current_device())
        row = np.asarray(row, np.int64)
        col = np.asarray(col, np.int64)
        n = Av.shape[0]
        if np.any(np.diff(row) < 0):
            raise ValueError('edges must be sorted by row')
        rowptr = np.concatenate([[0], np.cumsum(np.bincount(row, minlength=n))])
        key = row * n + col
        rkey = col * n + row
        order = np.argsort(key)
        pos = np.searchsorted(key[order], rkey)
        if np.any(pos >= key.shape[0]) or np.any(key[order][np.minimum(pos, key.shape[0] - 1)] != rkey):
            raise ValueError('the edge set is not symmetric')
        rev = order[pos]
        Av = np.asarray(Av, np.float64)
        Ae = np.asarray(Ae, np.float64)
        pts = np.asarray(points, np.float64)
        vfeat = np.stack([np.log(Av / Av.mean()),
                          np.asarray(logk_vertex, np.float64).reshape(-1)], 1).astype(np.float32)
        efeat = np.stack([Ae / np.sqrt(Av[row] * Av[col]),
                          (pts[col, 0] - pts[row, 0]) * n_grid,
                          (pts[col, 1] - pts[row, 1]) * n_grid], 1).astype(np.float32)
        i32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(dev)
        self.n, self.nnz = n, row.shape[0]
        self.rowptr, self.col, self.rev = i32(rowptr), i32(col), i32(rev)
        self.vfeat = torch.from_numpy(np.ascontiguousarray(vfeat)).to(dev)
        self.efeat = torch.from_numpy(np.ascontiguousarray(efeat)).to(dev)
        self.d = torch.from_numpy(Av).to(dev)
        self.device = dev
        self.slab = EdgeSlab(rowptr, col, efeat, n, dev) if use_slab else None


class GnnBuffers:
    """Inference buffers of the network on a mesh of n vertices: two
    (n, 32) bfloat16 activation and message buffers, ||r|| and the norm's
    reduction workspace."""

    def __init__(self, n, device):
        self.h = torch.empty((2, n, H), dtype=torch.bfloat16, device=device)
        self.Q = torch.empty((2, n, H), dtype=torch.bfloat16, device=device)
        self.rnorm = torch.empty(1, dtype=torch.float64, device=device)
        probe = torch.empty((n, 1), dtype=torch.float64, device=device)
        self.norm_ws = torch.empty(ops.aggregate_ws_bytes(probe), dtype=torch.uint8,
                                   device=device)


class LearnedPreconditioner:
    _bindings = 0

    def __init__(self, L=3, params: Optional[Dict[str, torch.Tensor]] = None,
                 device=None):
        self.binding = 0
        dev = device or torch.device('cuda', torch.cuda.current_device())
        npar, npk = ctypes.c_int64(), ctypes.c_int64()
        nat.call('esr_gnn_sizes', L, ctypes.byref(npar), ctypes.byref(npk))
        self.L, self.device = L, dev
        if params is None:
            params = init_params(L)
        flat = flatten(params, L)
        assert flat.numel() == npar.value, (flat.numel(), npar.value)
        self.params = flat.to(dev).contiguous()
        self.packed = torch.empty(npk.value, dtype=torch.bfloat16, device=dev)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grad = torch.zeros_like(self.params)
        self.step_count = 0
        self.graph: Optional[MeshGraph] = None
        self.pack()

    def pack(self):
        nat.call('esr_gnn_pack', self.L, self.params.data_ptr(), self.packed.data_ptr(),
                 nat.stream_ptr())

    def named(self):
        return unflatten(self.params.detach().cpu(), self.L)

    # -- binding to a mesh --------------------------------------------------
    def bind(self, graph: MeshGraph, train=False):
        LearnedPreconditioner._bindings += 1
        self.binding = LearnedPreconditioner._bindings   # buffers below are new
        self.graph = graph
        n, L = graph.n, self.L
        nh = (L + 1) if train else 2
        nq = L if train else 2
        self.h = torch.empty((nh, n, H), dtype=torch.bfloat16, device=self.device)
        self.Q = torch.empty((nq, n, H), dtype=torch.bfloat16, device=self.device)
        self.a = torch.empty((L, n, H), dtype=torch.float32, device=self.device) if train else None
        self.rnorm = torch.empty(1, dtype=torch.float64, device=self.device)
        probe = torch.empty((n, 1), dtype=torch.float64, device=self.device)
        self.norm_ws = torch.empty(ops.aggregate_ws_bytes(probe), dtype=torch.uint8,
                                   device=self.device)
        self.train = train
        if train:
            need = n * (8 * H + 3) + graph.nnz * H + 256 * 2 * H * H
            self.work = torch.empty(need, dtype=torch.float32, device=self.device)
        return self

    def apply(self, r: torch.Tensor, z: torch.Tensor):
        """z = M(r) (float64 device vectors of the bound mesh). Allocates
        nothing, so it can be captured in a HIP graph."""
        g = self.graph
        ops.aggregate('norm', r.reshape(-1, 1), out=self.rnorm, ws=self.norm_ws)
        _forward_call(self, g, r, self.rnorm, self.h, self.Q, self.a, 1 if self.train else 0, z)
        return z

    def forward_into(self, g, bufs: 'GnnBuffers', r: torch.Tensor, z: torch.Tensor):
        """z = M(r) on graph g (any object with n, rowptr, col, efeat, vfeat,
        d) with the inference buffers bufs -- several meshes at once, e.g. the
        levels of a multigrid hierarchy. Allocates nothing."""
        ops.aggregate('norm', r.reshape(-1, 1), out=bufs.rnorm, ws=bufs.norm_ws)
        _forward_call(self, g, r, bufs.rnorm, bufs.h, bufs.Q, None, 0, z)
        return z

    def backward(self, r: torch.Tensor, dz: torch.Tensor) -> torch.Tensor:
        """Gradient w.r.t. the flat parameters of sum(dz * M(r)); needs the
        preceding apply(r, .) in training mode."""
        if not self.train:
            raise RuntimeError('bind(graph, train=True) before backward()')
        g = self.graph
        nat.call('esr_gnn_backward', self.L, g.n, g.nnz, g.rowptr.data_ptr(),
                 g.col.data_ptr(), g.rev.data_ptr(), g.efeat.data_ptr(), g.vfeat.data_ptr(),
                 r.data_ptr(), g.d.data_ptr(), self.rnorm.data_ptr(), self.params.data_ptr(),
                 self.packed.data_ptr(), self.h.data_ptr(), self.Q.data_ptr(),
                 self.a.data_ptr(), dz.data_ptr(), self.grad.data_ptr(),
                 self.work.data_ptr(), self.work.numel(), nat.stream_ptr())
        return self.grad

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, clip=0.0):
        self.step_count += 1
        nat.call('esr_gnn_adam', self.params.numel(), self.params.data_ptr(),
                 self.grad.data_ptr(), self.m.data_ptr(), self.v.data_ptr(), lr, b1, b2,
                 eps, self.step_count, clip, nat.stream_ptr())
        self.pack()

    def __call__(self, r):
        z = torch.empty_like(r)
        return self.apply(r, z)


class GnnApply(torch.autograd.Function):
    """z = GnnApply.apply(params, r, pre): forward and backward run the
    hand-written kernels; params must be pre.params (the flat tensor)."""

    @staticmethod
    def forward(ctx, params, r, pre: LearnedPreconditioner):
        z = torch.empty_like(r)
        pre.apply(r, z)
        ctx.pre = pre
        ctx.save_for_backward(r)
        return z

    @staticmethod
    def backward(ctx, dz):
        (r,) = ctx.saved_tensors
        g = ctx.pre.backward(r, dz.contiguous().to(torch.float64))
        return g.clone(), None, None
