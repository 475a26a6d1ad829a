"""Multigrid kernels against the oracle (oracle/multigrid.py): the coarse
pattern and transfers bit-exact, Galerkin operators and V-cycles within
1e-12, and multigrid-preconditioned flexible CG through esr.compile."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import easier_amd as esr
from easier_amd import _native as nat
from easier_amd.numeric.preconditioner import multigrid as mg
from easier_amd.numeric.preconditioner.learned import LearnedPreconditioner
from easier_amd.numeric.solver import CG, FlexibleCG
from easier_amd.workloads import poisson as wp
from oracle import fem as ofem
from oracle import gnn
from oracle import mesh as omesh
from oracle import multigrid as omg

pytestmark = pytest.mark.gpu


def _problem(n, seed, permute=True):
    """Variable-coefficient P1 operator on tri2(n) in a random vertex order:
    (scipy A in lexicographic order, fpos, device CSR in the permuted order)."""
    m = omesh.tri2(n)
    rng = np.random.default_rng(seed)
    s = ofem.assemble(m, np.exp(rng.standard_normal(m['cells'].shape[0])))
    nv = s['Av'].size
    A = sp.csr_matrix((s['Ae'], (s['row'], s['col'])), shape=(nv, nv)) + sp.diags(s['Av'])
    fpos = rng.permutation(nv) if permute else np.arange(nv)   # new index -> lex
    inv = np.empty_like(fpos)
    inv[fpos] = np.arange(nv)
    # the assembly's pattern (boundary couplings stored as zeros), renumbered
    rows_new = inv[s['row']]
    cols_new = inv[s['col']]
    order = np.lexsort((cols_new, rows_new))
    rowptr = np.concatenate([[0], np.cumsum(np.bincount(rows_new, minlength=nv))])
    dev = torch.device('cuda', 0)
    i32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(dev)
    f64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)
    csr = (i32(rowptr), i32(cols_new[order]), f64(s['Ae'][order]), f64(s['Av'][fpos]))
    return A, fpos, csr, s


@pytest.mark.parametrize('nc', [1, 2, 5, 16])
def test_pattern_kernel(gpu, nc):
    W = nc + 1
    rowptr = torch.empty(W * W + 1, dtype=torch.int32, device=gpu)
    col = torch.empty((W - 1) * (6 * W - 2), dtype=torch.int32, device=gpu)
    nat.call('esr_mg_pattern', nc, rowptr.data_ptr(), col.data_ptr(), nat.stream_ptr())
    want_ptr, want_col = omg.pattern(nc)
    np.testing.assert_array_equal(rowptr.cpu().numpy(), want_ptr)
    np.testing.assert_array_equal(col.cpu().numpy(), want_col)


@pytest.mark.parametrize('permute', [False, True])
def test_hierarchy_matches_oracle_galerkin(gpu, permute):
    n = 32
    A, fpos, csr, _ = _problem(n, 3, permute)
    M = mg.MultigridPreconditioner(coarsest=4).setup(*csr, n, fpos)
    ref = omg.hierarchy(A[fpos][:, fpos], n, fpos, coarsest=4)
    assert [L.n_grid for L in M.levels] == [L.n for L in ref] == [32, 16, 8, 4]
    for L, R in zip(M.levels[1:], ref[1:]):
        rowptr, col, Ae, Av = omg.linsys_csr(R.A, L.n_grid)
        np.testing.assert_array_equal(L.rowptr.cpu().numpy(), rowptr)
        np.testing.assert_array_equal(L.col.cpu().numpy(), col)
        scale = abs(R.A).max()
        np.testing.assert_allclose(L.Ae.cpu().numpy(), Ae, rtol=0, atol=1e-13 * scale)
        np.testing.assert_allclose(L.Av.cpu().numpy(), Av, rtol=0, atol=1e-13 * scale)
    np.testing.assert_allclose(M.levels[-1].Ainv.cpu().numpy(), ref[-1].Ainv, rtol=1e-10,
                               atol=1e-12 * abs(ref[-1].Ainv).max())


def test_transfers_bit_exact(gpu):
    n = 16
    A, fpos, csr, _ = _problem(n, 5)
    fidx = np.empty_like(fpos)
    fidx[fpos] = np.arange(fpos.size)
    rng = np.random.default_rng(0)
    t = rng.standard_normal(fpos.size)
    zc = rng.standard_normal((n // 2 + 1) ** 2)
    z = rng.standard_normal(fpos.size)
    # every device buffer stays referenced until its kernel has run (a
    # temporary's block would go back to the allocator before the launch)
    fidx_d = torch.from_numpy(fidx.astype(np.int32)).to(gpu)
    fpos_d = torch.from_numpy(fpos.astype(np.int32)).to(gpu)
    t_d = torch.from_numpy(t).to(gpu)
    zc_d = torch.from_numpy(zc).to(gpu)
    zd = torch.from_numpy(z.copy()).to(gpu)
    rc = torch.empty(zc.size, dtype=torch.float64, device=gpu)
    nat.call('esr_mg_restrict', nat.F64, n, fidx_d.data_ptr(), t_d.data_ptr(),
             rc.data_ptr(), nat.stream_ptr())
    np.testing.assert_array_equal(rc.cpu().numpy(), omg.restrict(t, n, fidx))
    par = torch.empty(2 * fpos.size, dtype=torch.int32, device=gpu)
    nat.call('esr_mg_parents', n, fpos_d.data_ptr(), par.data_ptr(), nat.stream_ptr())
    # the parent map: interior coarse parents (-1 for boundary ones), -2 as
    # the second entry of a vertex on a coarse vertex
    pm = par.cpu().numpy().reshape(-1, 2)
    Wf, nc = n + 1, n // 2
    for i in range(fpos.size):
        I, J = fpos[i] % Wf, fpos[i] // Wf
        want = [b * (nc + 1) + a if 0 < a < nc and 0 < b < nc else -1
                for a, b, _ in omg.parents(I, J)]
        if len(want) == 1:
            want.append(-2)
        assert list(pm[i]) == want, (i, I, J, pm[i], want)
    nat.call('esr_mg_prolong', nat.F64, n, par.data_ptr(), zc_d.data_ptr(),
             zd.data_ptr(), nat.stream_ptr())
    np.testing.assert_array_equal(zd.cpu().numpy(), z + omg.prolong_values(zc, n, fpos))
    # lexicographic fine level (no maps)
    nat.call('esr_mg_restrict', nat.F64, n, None, t_d.data_ptr(), rc.data_ptr(),
             nat.stream_ptr())
    np.testing.assert_array_equal(rc.cpu().numpy(), omg.restrict(t, n))


@pytest.mark.parametrize('permute', [False, True])
def test_vcycle_matches_oracle(gpu, permute):
    n = 64
    A, fpos, csr, _ = _problem(n, 11, permute)
    M = mg.MultigridPreconditioner(coarsest=8).setup(*csr, n, fpos)
    ref = omg.hierarchy(A[fpos][:, fpos], n, fpos, coarsest=8)
    r = np.random.default_rng(2).standard_normal(fpos.size)
    z = M(torch.from_numpy(r).to(gpu)).cpu().numpy()
    want = omg.vcycle(ref, r, 0.8)
    np.testing.assert_allclose(z, want, rtol=0, atol=1e-12 * abs(want).max())
    # replays of a captured cycle give the same bits
    from easier_amd.numeric.solver.fused import _Graph
    rd = torch.from_numpy(r).to(gpu)
    zd = torch.zeros_like(rd)
    g = _Graph(lambda sp_: M.apply(rd, zd))
    outs = []
    for _ in range(3):
        zd.zero_()
        g.replay()
        torch.cuda.synchronize()
        outs.append(zd.cpu().numpy())
    for o in outs:
        np.testing.assert_array_equal(o, z)


def test_untrained_learned_smoother_is_undamped_jacobi(gpu):
    """The untrained network is z = r / d, so the learned cycle equals the
    oracle's Jacobi cycle with w = 1."""
    n = 32
    A, fpos, csr, s = _problem(n, 13, False)
    pts = omesh.tri2(n)['points']
    logk = np.zeros(fpos.size)
    from easier_amd.numeric.preconditioner.learned import MeshGraph
    graph = MeshGraph(s['row'], s['col'], s['Av'], s['Ae'], pts, logk, n, gpu)
    S = LearnedPreconditioner(3, gnn.init_params(3, 0))
    ref = omg.hierarchy(A, n, None, coarsest=4)
    r = np.random.default_rng(3).standard_normal(fpos.size)
    # the network before and after the coarse-grid correction
    M = mg.MultigridPreconditioner(S, coarsest=4, learned_min_grid=4, learned_post=True)
    M.setup(*csr, n, None, logk, graph)
    assert all(L.graph is not None for L in M.levels[:-1])
    z = M(torch.from_numpy(r).to(gpu)).cpu().numpy()
    want = omg.vcycle(ref, r, 1.0)
    np.testing.assert_allclose(z, want, rtol=0, atol=1e-12 * abs(want).max())
    # the default: network before, damped Jacobi (w = 0.8) after
    M = mg.MultigridPreconditioner(S, coarsest=4, learned_min_grid=4)
    M.setup(*csr, n, None, logk, graph)
    z = M(torch.from_numpy(r).to(gpu)).cpu().numpy()
    want = omg.vcycle(ref, r, 0.8, omega_pre=1.0)
    np.testing.assert_allclose(z, want, rtol=0, atol=1e-12 * abs(want).max())


def _compiled(n, seed, M):
    mesh, sys_, _ = wp.varcoef_system(n, seed)
    logk = wp.gaussian_random_field(n, seed).reshape(-1)
    eqn = wp.PoissonP1(sys_)
    sol = FlexibleCG(eqn.A, eqn.b, eqn.x, M, dict(points=mesh['points'], logk=logk, n_grid=n))
    esr.compile([eqn, sol], 'hip')
    return eqn, sol, sys_


@pytest.mark.parametrize('smoother', ['jacobi', 'learned'])
def test_multigrid_fcg_through_compile(gpu, smoother):
    n = 128
    S = LearnedPreconditioner(3, gnn.init_params(3, 0)) if smoother == 'learned' else None
    M = mg.MultigridPreconditioner(S, omega=0.8, learned_min_grid=32)
    eqn, sol, sys_ = _compiled(n, 7, M)
    info = sol.solve(rtol=1e-8, maxiter=500)
    bn = float(np.linalg.norm(sys_['b']))
    nv = sys_['Av'].size
    A = sp.csr_matrix((sys_['Ae'], (sys_['row'], sys_['col'])), shape=(nv, nv)) + \
        sp.diags(sys_['Av'])
    x = eqn.x.collect().numpy()
    assert np.linalg.norm(sys_['b'] - A @ x) <= 1.01e-8 * bn
    assert info['fallback_at'] is None
    # the untrained network smooths like undamped Jacobi (z = r / d), which
    # leaves the checkerboard mode alone: convergence, but slow
    assert info['iters'] <= (40 if smoother == 'jacobi' else 400), info
    eqn2 = wp.PoissonP1(sys_)
    cg = CG(eqn2.A, eqn2.b, eqn2.x, eqn2.M)
    esr.compile([eqn2, cg], 'hip')
    info2 = cg.solve(rtol=1e-8, maxiter=20000)
    if smoother == 'jacobi':
        # Jacobi-PCG on the same system needs an order of magnitude more steps
        assert info2['iters'] > 5 * info['iters'], (info, info2)
    x2 = eqn2.x.collect().numpy()
    assert np.linalg.norm(x - x2) <= 1e-6 * np.linalg.norm(x2)


@pytest.mark.parametrize('cycle_dtype', [torch.float64, torch.float32])
def test_multigrid_fcg_sell_and_csr_give_the_same_bits(gpu, cycle_dtype):
    """The level sweeps and flexible-CG products on the operators' SELL-64
    form and on their CSR give the same iterates: same steps, same residual,
    same solution bits."""
    from easier_amd.numeric.solver import fused
    print(f'multigrid FCG SELL vs CSR, cycle {cycle_dtype}', flush=True)
    out = []
    old = fused.FusedCG.use_sell
    try:
        for use in (True, False):
            fused.FusedCG.use_sell = use
            M = mg.MultigridPreconditioner(None, omega=0.8, cycle_dtype=cycle_dtype)
            eqn, sol, _ = _compiled(64, 3, M)
            info = sol.solve(rtol=1e-10, maxiter=200)
            assert all((L.sell is not None) == (use and L.Ainv is None) for L in M.levels)
            assert (sol.easier_kernels.sell is not None) == use
            out.append((info['iters'], info['residual'], eqn.x.collect()))
    finally:
        fused.FusedCG.use_sell = old
    (ia, ra, xa), (ib, rb, xb) = out
    print(f'  steps {ia} / {ib}, residual {ra:.6e} / {rb:.6e}', flush=True)
    assert ia == ib and ra == rb
    assert torch.equal(xa, xb)


def test_float32_cycle_matches_oracle(gpu):
    """The cycle in float32 (float32 copies of the float64 Galerkin
    operators) against the oracle's float64 cycle: float32 rounding only."""
    n = 64
    A, fpos, csr, _ = _problem(n, 17, True)
    M = mg.MultigridPreconditioner(coarsest=8, cycle_dtype=torch.float32).setup(*csr, n, fpos)
    assert M.levels[0].z.dtype == torch.float32
    ref = omg.hierarchy(A[fpos][:, fpos], n, fpos, coarsest=8)
    r = np.random.default_rng(5).standard_normal(fpos.size)
    z = M(torch.from_numpy(r).to(gpu))
    assert z.dtype == torch.float64
    want = omg.vcycle(ref, r, 0.8)
    err = np.abs(z.cpu().numpy() - want).max() / np.abs(want).max()
    assert err <= 1e-5, err


def test_float32_cycle_fcg_reaches_float64_tolerance(gpu):
    """Mixed precision: flexible CG in float64 (dot products, iteration,
    true residual) with a float32 cycle reaches 1e-10 on the true residual,
    in about the steps of the float64 cycle."""
    n = 128
    its = {}
    for cd in (torch.float64, torch.float32):
        M = mg.MultigridPreconditioner(None, cycle_dtype=cd)
        eqn, sol, sys_ = _compiled(n, 21, M)
        info = sol.solve(rtol=1e-10, maxiter=500)
        nv = sys_['Av'].size
        A = sp.csr_matrix((sys_['Ae'], (sys_['row'], sys_['col'])), shape=(nv, nv)) + \
            sp.diags(sys_['Av'])
        x = eqn.x.collect().numpy()
        assert np.linalg.norm(sys_['b'] - A @ x) <= 1.01e-10 * np.linalg.norm(sys_['b'])
        assert info['fallback_at'] is None
        its[cd] = info['iters']
    assert its[torch.float32] <= its[torch.float64] + 3, its


def test_multigrid_on_jittered_mesh(gpu):
    """Interior vertices jittered by up to 0.2 h (the batched family's
    meshes): grid positions are recovered from the coordinates, P is the
    index-based interpolation (no longer exact P1 nesting), and flexible CG
    still converges in a few dozen steps on the true residual."""
    n = 128
    mesh, sys_, kap = wp.varcoef_system(n, 5, jitter=0.2)
    logk = wp.gaussian_random_field(n, 5).reshape(-1)
    eqn = wp.PoissonP1(sys_)
    sol = FlexibleCG(eqn.A, eqn.b, eqn.x, mg.MultigridPreconditioner(None),
                     dict(points=mesh['points'], logk=logk, n_grid=n))
    esr.compile([eqn, sol], 'hip')
    info = sol.solve(rtol=1e-8, maxiter=500)
    nv = sys_['Av'].size
    A = sp.csr_matrix((sys_['Ae'], (sys_['row'], sys_['col'])), shape=(nv, nv)) + \
        sp.diags(sys_['Av'])
    x = eqn.x.collect().numpy()
    assert np.linalg.norm(sys_['b'] - A @ x) <= 1.01e-8 * np.linalg.norm(sys_['b'])
    assert info['fallback_at'] is None and info['iters'] <= 40, info


def test_setup_rejects_meshes_it_cannot_coarsen(gpu):
    n = 100                      # 100 -> 50 -> 25: an odd grid of 25 <= 32, accepted
    A, fpos, csr, _ = _problem(n, 1, False)
    M = mg.MultigridPreconditioner(coarsest=16).setup(*csr, n, fpos)
    assert [L.n_grid for L in M.levels] == [100, 50, 25]
    n = 2 * 37                   # 74 -> 37: odd and above 32
    A, fpos, csr, _ = _problem(n, 1, False)
    with pytest.raises(ValueError, match='coarsest'):
        mg.MultigridPreconditioner(coarsest=16, coarsening='geometric').setup(*csr, n, fpos)
    with pytest.raises(ValueError, match='grid'):
        mg.MultigridPreconditioner(coarsening='geometric').setup(*csr, n + 1, fpos)
    # the default coarsens such operators by aggregation instead
    M = mg.MultigridPreconditioner(coarsest=16).setup(*csr, n, fpos)
    assert M.kind == 'aggregation' and M.levels[-1].nv <= 1024
