"""The learned preconditioner's kernels against the CPU reference network:
forward (MFMA, bfloat16 operands, float32 accumulation), backward (float32
kernels, deterministic reductions), the autograd.Function wrapper, and a
finite-difference check of the device gradient."""
import numpy as np
import pytest
import torch

from easier_amd.numeric.preconditioner.learned import (GnnApply, LearnedPreconditioner,
                                                       MeshGraph, flatten, unflatten)
from easier_amd.workloads import mesh as wmesh
from easier_amd.workloads import poisson as wp
from oracle import gnn

pytestmark = pytest.mark.gpu


def _setup(n=12, seed=0, L=3, kind='tri2'):
    # tri2: rows of at most 6 couplings; tri4: square corners have 8
    mesh = wmesh.tri2(n, jitter=0.2, seed=seed) if kind == 'tri2' else wmesh.tri4(n)
    rng = np.random.default_rng(seed)
    logk = rng.standard_normal(mesh['points'].shape[0])
    s = wp.p1_system(mesh, wp.kappa_cells(mesh, logk), None)
    vfeat, efeat = gnn.static_features(s['Av'], s['Ae'], s['row'], s['col'],
                                       mesh['points'], logk, n)
    static = (torch.from_numpy(s['Av']), vfeat, efeat, torch.from_numpy(s['row']),
              torch.from_numpy(s['col']))
    graph = MeshGraph(s['row'], s['col'], s['Av'], s['Ae'], mesh['points'], logk, n)
    p = gnn.init_params(L, seed)
    g = torch.Generator().manual_seed(seed + 7)
    p['w_out'] = torch.randn(gnn.H, generator=g) * 0.3
    p['b_out'] = torch.tensor(0.05)
    for k in p:
        if k.startswith('b') and k != 'b_out':
            p[k] = 0.05 * torch.randn(p[k].shape, generator=g)
    return s, static, graph, p


def _rel(a, b):
    return float((a - b).norm() / b.norm())


def test_untrained_network_is_exactly_jacobi(gpu):
    s, static, graph, _ = _setup()
    pre = LearnedPreconditioner(3, gnn.init_params(3, 0)).bind(graph)
    r = torch.randn(graph.n, dtype=torch.float64, generator=torch.Generator().manual_seed(1))
    z = pre(r.to(gpu)).cpu()
    assert torch.equal(z, r / torch.from_numpy(s['Av']))


@pytest.mark.parametrize('seed,kind', [(0, 'tri2'), (1, 'tri2'), (2, 'tri4')])
def test_forward_matches_reference(gpu, seed, kind):
    print(f'gnn forward seed={seed} mesh={kind}', flush=True)
    s, static, graph, p = _setup(seed=seed, kind=kind)
    pre = LearnedPreconditioner(3, p).bind(graph)
    r = torch.randn(graph.n, dtype=torch.float64,
                    generator=torch.Generator().manual_seed(seed + 3))
    z = pre(r.to(gpu)).cpu()
    d, vfeat, efeat, row, col = static
    zr = gnn.forward(p, r, d, vfeat, efeat, row, col, 3)
    jac = r / d
    # compare the network's contribution (z - r/d); bfloat16 rounding of the
    # hidden state can flip at ties, so the bound is relative, not bitwise
    assert _rel(z - jac, zr - jac) < 2e-2
    assert _rel(z, zr) < 2e-3


@pytest.mark.parametrize('n,kind', [(12, 'tri2'), (12, 'tri4'), (37, 'tri2')])
def test_slab_forward_is_bit_identical_to_csr(gpu, n, kind):
    """The forward on the slot-major edge slab (esr_gnn_forward_slab) gives
    the bits of the forward on the CSR: the same edges folded in the same
    order; tri4 has vertices with 8 edges (the loop past the 6 unrolled
    slots), n = 37 a partial last group of 16 vertices."""
    print(f'gnn slab vs csr n={n} mesh={kind}', flush=True)
    s, static, graph, p = _setup(n=n, kind=kind)
    assert graph.slab is not None and graph.slab.slots >= graph.nnz
    pre = LearnedPreconditioner(3, p).bind(graph)
    r = torch.randn(graph.n, dtype=torch.float64,
                    generator=torch.Generator().manual_seed(5)).to(gpu)
    z_slab = pre(r).cpu()
    slab, graph.slab = graph.slab, None
    z_csr = pre(r).cpu()
    graph.slab = slab
    assert torch.equal(z_slab, z_csr)


@pytest.mark.parametrize('kind', ['tri2', 'tri4'])
def test_backward_matches_reference_autograd(gpu, kind):
    print(f'gnn backward mesh={kind}', flush=True)
    s, static, graph, p = _setup(seed=2, kind=kind)
    L = 3
    pre = LearnedPreconditioner(L, p).bind(graph, train=True)
    gen = torch.Generator().manual_seed(5)
    r = torch.randn(graph.n, dtype=torch.float64, generator=gen)
    dz = torch.randn(graph.n, dtype=torch.float64, generator=gen)
    pre.apply(r.to(gpu), torch.empty(graph.n, dtype=torch.float64, device=gpu))
    grad = unflatten(pre.backward(r.to(gpu), dz.to(gpu)).cpu(), L)
    d, vfeat, efeat, row, col = static
    pp = {k: v.clone().requires_grad_(True) for k, v in p.items()}
    z = gnn.forward(pp, r, d, vfeat, efeat, row, col, L)
    ref = torch.autograd.grad((dz * z).sum(), list(pp.values()))
    for (k, _), gr in zip(pp.items(), ref):
        err = float((grad[k] - gr.float()).norm()) / max(float(gr.norm()), 1e-12)
        assert err < 5e-2, (k, err)
    # deterministic: same bits on a second backward
    g1 = pre.grad.clone()
    pre.apply(r.to(gpu), torch.empty(graph.n, dtype=torch.float64, device=gpu))
    pre.backward(r.to(gpu), dz.to(gpu))
    assert torch.equal(g1, pre.grad)


def test_autograd_function_and_finite_differences(gpu):
    s, static, graph, p = _setup(n=16, seed=3)
    L = 3
    pre = LearnedPreconditioner(L, p).bind(graph, train=True)
    gen = torch.Generator().manual_seed(6)
    r = torch.randn(graph.n, dtype=torch.float64, generator=gen).to(gpu)
    c = torch.randn(graph.n, dtype=torch.float64, generator=gen).to(gpu)
    params = pre.params.requires_grad_(True)
    z = GnnApply.apply(params, r, pre)
    (g_auto,) = torch.autograd.grad((c * z).sum(), params)
    params.requires_grad_(False)
    # directional finite difference of the device forward along the gradient
    v = g_auto / g_auto.norm()
    base = pre.params.clone()
    vals = []
    eps = 2e-2
    for sgn in (1.0, -1.0):
        pre.params.copy_(base + sgn * eps * v)
        pre.pack()
        vals.append(float((c * pre(r)).sum()))
    pre.params.copy_(base)
    pre.pack()
    fd = (vals[0] - vals[1]) / (2 * eps)
    an = float((g_auto * v).sum())
    assert abs(fd - an) <= 0.1 * abs(an), (fd, an)


def test_forward_timing_command(gpu):
    """`python -m easier_amd.workloads.learned forward` on a small mesh."""
    from easier_amd.workloads import learned as wl
    pre = LearnedPreconditioner(3, gnn.init_params(3, 0))
    res = wl.time_forward(pre, 32, seed=2, reps=3)
    assert res['vertices'] == 33 * 33 and res['us_per_application'] > 0
