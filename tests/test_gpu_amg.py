"""Algebraic multigrid (smoothed aggregation) on the GPU: one V-cycle against
the oracle's restatement (oracle/amg.py) in float64 and float32, and flexible
CG preconditioned by it through esr.compile on meshes geometric coarsening
cannot take -- odd grids, the reference's 4-triangle mesh (P1 and the
tutorial's finite-volume operator) -- with the true residual checked on the
host and the solution against a direct solve."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

import easier_amd as esr
from easier_amd.numeric.preconditioner import multigrid as mg
from easier_amd.numeric.preconditioner.learned import LearnedPreconditioner
from easier_amd.numeric.solver import FlexibleCG
from easier_amd.numeric.solver.fused import FusedCG
from easier_amd.workloads import mesh as wmesh
from easier_amd.workloads import poisson as wp
from easier_amd.workloads import poisson_fv as wfv
from oracle import amg as oamg
from oracle import fv_poisson as ofv
from oracle import gnn

pytestmark = pytest.mark.gpu


def _scipy(s):
    nv = s['Av'].size
    return sp.csr_matrix((s['Ae'], (s['row'], s['col'])), shape=(nv, nv)) + sp.diags(s['Av'])


def _device_csr(s, dev):
    nv = s['Av'].size
    rowptr = np.concatenate([[0], np.cumsum(np.bincount(s['row'], minlength=nv))])
    i32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(dev)
    f64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)
    return i32(rowptr), i32(s['col']), f64(s['Ae']), f64(s['Av'])


@pytest.mark.parametrize('cycle_dtype', [torch.float64, torch.float32])
@pytest.mark.parametrize('sell', [True, False])
def test_aggregation_cycle_matches_oracle(gpu, cycle_dtype, sell, monkeypatch):
    print(f'AMG cycle tri2 n=37 jittered {cycle_dtype} sell={sell}', flush=True)
    monkeypatch.setattr(FusedCG, 'use_sell', sell)
    _, s, _ = wp.varcoef_system(37, 4, jitter=0.2)
    A = _scipy(s)
    M = mg.MultigridPreconditioner(None, coarsening='aggregation', coarse_rows=64,
                                   cycle_dtype=cycle_dtype).setup(*_device_csr(s, gpu))
    ref = oamg.hierarchy(A, coarse_rows=64)
    assert M.kind == 'aggregation'
    assert [L.nv for L in M.levels] == [L.A.shape[0] for L in ref]
    assert len(M.levels) >= 3
    # the finest level reads its SELL-64 form when enabled; coarse algebraic
    # levels spread rows over lanes of the CSR kernels
    assert (M.levels[0].sell is not None) == sell
    assert all(L.sell is None for L in M.levels[1:])
    assert all(L.lanes_A > 1 for L in M.levels[1:-1])
    for r_seed in (5, 6):
        r = np.random.default_rng(r_seed).standard_normal(A.shape[0])
        z = M(torch.from_numpy(r).to(gpu))
        want = oamg.vcycle(ref, r, 0.8)
        err = np.abs(z.cpu().numpy() - want).max() / np.abs(want).max()
        assert err <= (1e-12 if cycle_dtype == torch.float64 else 1e-5), err


def _solve(eqn, M, features=None, rtol=1e-8):
    sol = FlexibleCG(eqn.A, eqn.b, eqn.x, M, features)
    esr.compile([eqn, sol], 'hip')
    return sol.solve(rtol=rtol, maxiter=500)


def _check(A, b, x, info, rtol=1e-8, steps=30):
    b = np.asarray(b)
    assert np.linalg.norm(b - A @ x) <= 1.01 * rtol * np.linalg.norm(b)
    assert info['fallback_at'] is None and info['iters'] <= steps, info
    x_d = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - x_d) <= 1e-6 * np.linalg.norm(x_d)


@pytest.mark.parametrize('case', ['tri2-37-jittered', 'tri2-101', 'tri4-40'])
def test_aggregation_fcg_where_geometric_coarsening_does_not_apply(gpu, case):
    print(f'AMG flexible CG {case}', flush=True)
    if case == 'tri4-40':
        m = wmesh.tri4(40)
        rng = np.random.default_rng(9)
        s = wp.p1_system(m, np.exp(rng.standard_normal(m['cells'].shape[0])),
                         wp.rhs_uniform(m['points'].shape[0], 10))
        features = None
    else:
        n = 37 if case.startswith('tri2-37') else 101
        m, s, _ = wp.varcoef_system(n, 11, jitter=0.2 if 'jitter' in case else 0.0)
        features = dict(points=m['points'], n_grid=n)   # geometric: odd grid above 32
    eqn = wp.PoissonP1(s)
    M = mg.MultigridPreconditioner(None)
    info = _solve(eqn, M, features)
    print(f'  {len(M.levels)} levels {[L.nv for L in M.levels]}, {info["iters"]} steps',
          flush=True)
    assert M.kind == 'aggregation'
    _check(_scipy(s), s['b'], eqn.x.collect().numpy(), info)


def test_aggregation_fcg_on_the_reference_fv_poisson_operator(gpu):
    """The tutorial's finite-volume operator on the reference's 4-triangle
    mesh (n = 100, 40,000 cells) is negative definite (Ac < 0 < Af); the
    same solution solves the SPD system -A x = -b."""
    print('AMG flexible CG FV Poisson tri4 n=100 (negated)', flush=True)
    mesh = wmesh.tri4(100)
    ref = ofv.assemble(mesh)
    neg = {k: -ref[k] for k in ('Ac', 'Af', 'b')}
    eqn = wfv.FVPoisson(mesh, neg)
    M = mg.MultigridPreconditioner(None)
    info = _solve(eqn, M)
    nc = mesh['cells'].shape[0]
    A = sp.csr_matrix((neg['Af'].numpy(), (mesh['src'], mesh['dst'])), shape=(nc, nc)) + \
        sp.diags(neg['Ac'].numpy())
    print(f'  {len(M.levels)} levels, {info["iters"]} steps', flush=True)
    assert M.kind == 'aggregation'
    _check(A, neg['b'].numpy(), eqn.x.collect().numpy(), info)


def test_learned_smoother_on_an_aggregation_hierarchy(gpu):
    """The network smooths the finest level (edge vectors in units of 1/37);
    the algebraic coarse levels use damped Jacobi."""
    print('AMG + learned smoother tri2 n=37 jittered', flush=True)
    n = 37
    m, s, _ = wp.varcoef_system(n, 12, jitter=0.2)
    logk = wp.gaussian_random_field(n, 12).reshape(-1)
    eqn = wp.PoissonP1(s)
    S = LearnedPreconditioner(3, gnn.init_params(3, 0))
    M = mg.MultigridPreconditioner(S, coarsening='auto')
    info = _solve(eqn, M, dict(points=m['points'], logk=logk, n_grid=n))
    assert M.kind == 'aggregation'
    assert M.levels[0].graph is not None
    assert all(L.graph is None for L in M.levels[1:])
    _check(_scipy(s), s['b'], eqn.x.collect().numpy(), info)


def test_auto_takes_geometric_coarsening_where_it_applies(gpu):
    print('multigrid auto vs aggregation tri2 n=64', flush=True)
    n = 64
    steps = {}
    for kind in ('auto', 'aggregation'):
        m, s, _ = wp.varcoef_system(n, 13)
        eqn = wp.PoissonP1(s)
        M = mg.MultigridPreconditioner(None, coarsening=kind)
        info = _solve(eqn, M, dict(points=m['points'], n_grid=n))
        assert M.kind == ('geometric' if kind == 'auto' else 'aggregation')
        _check(_scipy(s), s['b'], eqn.x.collect().numpy(), info)
        steps[kind] = info['iters']
    assert steps['aggregation'] <= steps['auto'] + 10, steps


def test_float32_aggregation_cycle_reaches_float64_tolerance(gpu):
    print('AMG float32 cycle tri2 n=75 jittered, rtol 1e-10', flush=True)
    n = 75
    its = {}
    for cd in (torch.float64, torch.float32):
        m, s, _ = wp.varcoef_system(n, 14, jitter=0.2)
        eqn = wp.PoissonP1(s)
        M = mg.MultigridPreconditioner(None, coarsening='aggregation', cycle_dtype=cd)
        info = _solve(eqn, M, rtol=1e-10)
        _check(_scipy(s), s['b'], eqn.x.collect().numpy(), info, rtol=1e-10, steps=40)
        its[cd] = info['iters']
    assert its[torch.float32] <= its[torch.float64] + 3, its


def test_config3_family_with_the_trained_smoother(gpu):
    """The learned preconditioner on the batched configuration's problem
    family: sides n drawn uniformly from [32, 512] (the batched benchmark's
    draw), interior vertices jittered by up to 0.2 h, a seeded log-kappa
    field. Most such n are not m 2^k with m <= 32, so the multigrid takes
    the aggregation hierarchy there; with and without the trained network
    flexible CG reaches 1e-8 on the true residual in a few dozen steps."""
    import os
    from easier_amd.numeric.preconditioner import learned as lp
    from easier_amd.workloads import learned as wl
    S = lp.load(os.path.join(os.path.dirname(wl.__file__), 'gnn_smoother_l3_h32.npz'))
    rng = np.random.default_rng(0)
    sides = [int(rng.integers(32, 513)) for _ in range(4)]
    kinds = set()
    for k, n in enumerate(sides):
        m, s, _ = wp.varcoef_system(n, 100 + k, jitter=0.2)
        logk = wp.gaussian_random_field(n, 100 + k).reshape(-1)
        for smoother in (None, S):
            print(f'config-3 family n={n} smoother={smoother is not None}', flush=True)
            eqn = wp.PoissonP1(s)
            M = mg.MultigridPreconditioner(smoother)
            info = _solve(eqn, M, dict(points=m['points'], logk=logk, n_grid=n))
            geometric = M.geometric_unfit(s['Av'].size, n) is None
            assert M.kind == ('geometric' if geometric else 'aggregation')
            kinds.add(M.kind)
            print(f'  {M.kind}, {info["iters"]} steps', flush=True)
            _check(_scipy(s), s['b'], eqn.x.collect().numpy(), info, steps=40)
    assert 'aggregation' in kinds


def test_operator_without_strong_couplings_is_smoothed_only(gpu):
    """theta above every coupling: nothing aggregates, the level is too
    large for the dense solve, and the cycle is two damped-Jacobi sweeps
    (weight 0.8 * 2 / rho); flexible CG still converges."""
    print('AMG smoother-only level tri2 n=37', flush=True)
    _, s, _ = wp.varcoef_system(37, 15, jitter=0.2)
    A = _scipy(s)
    M = mg.MultigridPreconditioner(None, coarsening='aggregation', theta=0.999,
                                   coarse_rows=64).setup(*_device_csr(s, gpu))
    assert len(M.levels) == 1 and M.levels[0].Ainv is None
    d = A.diagonal()
    rho = np.max(np.abs(A).sum(axis=1).A1 / d)
    w = 0.8 * 2.0 / rho
    r = np.random.default_rng(3).standard_normal(A.shape[0])
    z = M(torch.from_numpy(r).to(gpu)).cpu().numpy()
    z0 = w * r / d
    want = z0 + w * (r - A @ z0) / d
    assert np.abs(z - want).max() <= 1e-12 * np.abs(want).max()
    eqn = wp.PoissonP1(s)
    sol = FlexibleCG(eqn.A, eqn.b, eqn.x, mg.MultigridPreconditioner(
        None, coarsening='aggregation', theta=0.999, coarse_rows=64))
    esr.compile([eqn, sol], 'hip')
    info = sol.solve(rtol=1e-8, maxiter=2000)
    _check(A, s['b'], eqn.x.collect().numpy(), info, steps=2000)
