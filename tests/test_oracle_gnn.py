"""The learned-preconditioner reference network: its autograd gradients
against central finite differences (float64, no bfloat16 rounding), and the
properties the solver relies on."""
import numpy as np
import pytest
import torch

from easier_amd.workloads import mesh as wmesh
from easier_amd.workloads import poisson as wp
from oracle import gnn


def _problem(n=4, seed=0):
    mesh = wmesh.tri2(n, jitter=0.2, seed=seed)
    rng = np.random.default_rng(seed)
    logk = rng.standard_normal(mesh['points'].shape[0])
    kap = wp.kappa_cells(mesh, logk)
    s = wp.p1_system(mesh, kap, None)
    vfeat, efeat = gnn.static_features(s['Av'], s['Ae'], s['row'], s['col'],
                                       mesh['points'], logk, n)
    static = (torch.from_numpy(s['Av']), vfeat, efeat, torch.from_numpy(s['row']),
              torch.from_numpy(s['col']))
    return s, static


def _random_params(L, seed):
    p = gnn.init_params(L, seed)
    g = torch.Generator().manual_seed(seed + 100)
    p['w_out'] = torch.randn(gnn.H, generator=g) * 0.3
    p['b_out'] = torch.tensor(0.1)
    for k in p:
        if k.startswith('b'):
            p[k] = p[k] + 0.05 * torch.randn(p[k].shape, generator=g)
    return p


def test_gradients_match_finite_differences():
    s, static = _problem()
    L = 2
    p = {k: v.double().requires_grad_(True) for k, v in _random_params(L, 1).items()}
    n = s['Av'].shape[0]
    g = torch.Generator().manual_seed(3)
    r = torch.randn(n, generator=g, dtype=torch.float64)
    c = torch.randn(n, generator=g, dtype=torch.float64)
    d, vfeat, efeat, row, col = static

    def f(pp):
        return (c * gnn.forward(pp, r, d, vfeat, efeat, row, col, L, bf16=False,
                                dtype=torch.float64)).sum()
    val = f(p)
    grads = torch.autograd.grad(val, list(p.values()))
    rng = np.random.default_rng(0)
    checked = 0
    for (name, t), gr in zip(p.items(), grads):
        flat = t.detach().reshape(-1)
        for idx in rng.choice(flat.numel(), size=min(3, flat.numel()), replace=False):
            eps = 1e-6
            pp = {k: v.detach().clone() for k, v in p.items()}
            pp[name].reshape(-1)[idx] += eps
            up = f(pp)
            pp[name].reshape(-1)[idx] -= 2 * eps
            dn = f(pp)
            fd = float(up - dn) / (2 * eps)
            an = float(gr.reshape(-1)[idx])
            assert abs(fd - an) <= 1e-5 * max(1.0, abs(an)), (name, idx, fd, an)
            checked += 1
    assert checked > 30


def test_untrained_network_is_jacobi_and_scale_equivariant():
    s, static = _problem(seed=2)
    d, vfeat, efeat, row, col = static
    n = s['Av'].shape[0]
    r = torch.randn(n, generator=torch.Generator().manual_seed(4), dtype=torch.float64)
    p0 = gnn.init_params(3, 0)
    z = gnn.forward(p0, r, d, vfeat, efeat, row, col, 3)
    assert torch.equal(z, r / d)
    p = _random_params(3, 5)
    z1 = gnn.forward(p, r, d, vfeat, efeat, row, col, 3, bf16=False, dtype=torch.float64)
    z2 = gnn.forward(p, 4.0 * r, d, vfeat, efeat, row, col, 3, bf16=False,
                     dtype=torch.float64)
    torch.testing.assert_close(z2, 4.0 * z1, rtol=1e-12, atol=0)
