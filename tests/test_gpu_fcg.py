"""Flexible CG with the learned preconditioner: convergence on the true
float64 residual, the Jacobi fallback, and that training lowers the loss
below the untrained (Jacobi) network's."""
import numpy as np
import pytest
import torch

import easier_amd as esr
from easier_amd.numeric.preconditioner.learned import LearnedPreconditioner
from easier_amd.numeric.solver import CG, FlexibleCG
from easier_amd.workloads import learned as wl
from easier_amd.workloads import poisson as wp
from oracle import gnn
from oracle import solvers as osol

pytestmark = pytest.mark.gpu


def _program(n, seed, M):
    mesh, sys_, kap = wp.varcoef_system(n, seed)
    logk = wp.gaussian_random_field(n, seed).reshape(-1)
    eqn = wp.PoissonP1(sys_)
    sol = FlexibleCG(eqn.A, eqn.b, eqn.x, M,
                     dict(points=mesh['points'], logk=logk, n_grid=n))
    esr.compile([eqn, sol], 'hip')
    return eqn, sol, sys_


def _true_residual(sys_, x):
    A = osol.Linsys(torch.from_numpy(sys_['Av']), torch.from_numpy(sys_['Ae']),
                    torch.from_numpy(sys_['col']), torch.from_numpy(sys_['row']),
                    sys_['Av'].shape[0])
    return float((torch.from_numpy(sys_['b']) - A(x)).norm())


def test_untrained_network_behaves_like_jacobi_pcg(gpu):
    M = LearnedPreconditioner(3, gnn.init_params(3, 0))
    eqn, sol, sys_ = _program(24, 1, M)
    info = sol.solve(rtol=1e-8, maxiter=5000)
    bn = float(np.linalg.norm(sys_['b']))
    assert info['residual'] <= 1e-8 * bn and info['fallback_at'] is None
    assert _true_residual(sys_, eqn.x.collect()) <= 1.01e-8 * bn
    # against the fused Jacobi-PCG on the same system
    eqn2 = wp.PoissonP1(sys_)
    cg = CG(eqn2.A, eqn2.b, eqn2.x, eqn2.M)
    esr.compile([eqn2, cg], 'hip')
    info2 = cg.solve(rtol=1e-8, maxiter=5000)
    assert abs(info['iters'] - (info2['iters'] + 1)) <= 3
    x1, x2 = eqn.x.collect(), eqn2.x.collect()
    assert float((x1 - x2).norm()) <= 1e-5 * float(x2.norm())


def test_random_network_converges_and_fallback_path(gpu):
    p = gnn.init_params(3, 4)
    p['w_out'] = torch.randn(gnn.H, generator=torch.Generator().manual_seed(2)) * 0.5
    M = LearnedPreconditioner(3, p)
    eqn, sol, sys_ = _program(20, 2, M)
    info = sol.solve(rtol=1e-8, maxiter=5000)
    bn = float(np.linalg.norm(sys_['b']))
    assert info['residual'] <= 1e-8 * bn
    assert _true_residual(sys_, eqn.x.collect()) <= 1.01e-8 * bn
    eqn.x.data.zero_()
    info = sol.solve(rtol=1e-8, maxiter=5000, window=-1)   # stagnation at once
    assert info['fallback_at'] is not None
    assert info['residual'] <= 1e-8 * bn
    assert _true_residual(sys_, eqn.x.collect()) <= 1.01e-8 * bn


@pytest.mark.parametrize('kind', ['energy', 'error', 'twogrid', 'twogrid_post'])
def test_pair_loss_gradient_matches_finite_differences(gpu, kind):
    prob = wl.Problem(12, 3)
    r, e, en2, re = wl.generate_pairs(prob, n_rhs=1, seed=0)[3]
    # r.e is the A-norm of e: A e = r
    assert re == pytest.approx(float((wl.spmv(prob, e.double()) * e.double()).sum()),
                               rel=1e-5)
    g = torch.Generator().manual_seed(1)
    z = (r / prob.Av) + 0.1 * torch.randn(prob.n, generator=g, dtype=torch.float64).to(gpu)
    v = torch.randn(prob.n, generator=g, dtype=torch.float64).to(gpu)
    _, dz = wl.pair_loss(prob, r, e, en2, re, z, kind)
    h = 1e-6 * float(z.norm()) / float(v.norm())
    lp_ = wl.pair_loss(prob, r, e, en2, re, z + h * v, kind)[0]
    lm_ = wl.pair_loss(prob, r, e, en2, re, z - h * v, kind)[0]
    fd = (lp_ - lm_) / (2 * h)
    assert float((dz * v).sum()) == pytest.approx(fd, rel=1e-5, abs=1e-12)
    if kind.startswith('twogrid'):
        # smoothing with the exact error leaves nothing; no smoothing leaves
        # the part of e outside the coarse space, a fraction in (0, 1)
        ed = e.double()
        assert abs(wl.pair_loss(prob, r, e, en2, re, ed, kind)[0]) < 1e-5
        l0 = wl.pair_loss(prob, r, e, en2, re, torch.zeros_like(ed), kind)[0]
        assert 0.0 < l0 < 1.0


def test_training_beats_jacobi_on_the_training_loss(gpu):
    probs = [wl.Problem(20, s) for s in (10, 11)]
    data = [(pr, wl.generate_pairs(pr, n_rhs=4, seed=s)) for s, pr in enumerate(probs)]
    assert sum(len(p) for _, p in data) >= 20
    base = wl.mean_loss(None, data)
    pre = LearnedPreconditioner(3, gnn.init_params(3, 0))
    # the untrained network is Jacobi
    assert wl.mean_loss(pre, data) == pytest.approx(base, rel=1e-6)
    wl.train(pre, data, epochs=6, lr=2e-3)
    trained = wl.mean_loss(pre, data)
    assert trained < 0.98 * base, (trained, base)


def test_trained_smoother_beats_damped_jacobi_on_the_two_grid_error(gpu):
    """What training gains: on the post-smoothed two-grid cycle the network
    is trained for (network pre-smoothing, exact coarse correction, damped
    Jacobi post-smoothing), the trained network leaves less A-norm error than
    damped Jacobi (w = 0.8) and than the untrained network (w = 1), on pairs
    from the plain solver and random errors -- and on held-out problems."""
    train_probs = [wl.Problem(32, s) for s in (30, 31)]
    held_out = [wl.Problem(32, 40)]

    def pairs(pr, s):
        return wl.generate_pairs(pr, n_rhs=3, seed=s) + \
            wl.generate_pairs_random(pr, 24, seed=100 + s)
    data = [(pr, pairs(pr, s)) for s, pr in enumerate(train_probs)]
    test = [(pr, pairs(pr, 50 + s)) for s, pr in enumerate(held_out)]
    kind = 'twogrid_post'
    jac1 = wl.mean_loss(None, data, kind)
    jac08 = wl.mean_loss(None, data, kind, omega=0.8)
    assert jac08 < jac1
    pre = LearnedPreconditioner(3, gnn.init_params(3, 0))
    assert wl.mean_loss(pre, data, kind) == pytest.approx(jac1, rel=1e-6)
    wl.train(pre, data, epochs=6, lr=2e-3, loss=kind)
    trained = wl.mean_loss(pre, data, kind)
    assert trained < 0.7 * jac08, (trained, jac08, jac1)
    held = wl.mean_loss(pre, test, kind)
    assert held < wl.mean_loss(None, test, kind, omega=0.8), held
