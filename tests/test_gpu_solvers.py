"""Batched Jacobi-PCG on the GPU against per-instance oracle solves."""
import numpy as np
import pytest
import torch

from easier_amd.numeric.solver.batched import BatchedCG
from easier_amd.workloads import mesh as wmesh
from easier_amd.workloads import poisson as wp
from oracle import solvers as osol

pytestmark = pytest.mark.gpu


def _oracle_solve(s, rtol):
    A = osol.Linsys(torch.from_numpy(s['Av']), torch.from_numpy(s['Ae']),
                    torch.from_numpy(s['col']), torch.from_numpy(s['row']),
                    s['Av'].shape[0])
    b = torch.from_numpy(s['b'])
    x = torch.zeros_like(b)
    info, x = osol.cg(A, b, x, M=lambda r: r / A.Av, rtol=rtol, maxiter=20000)
    return info, x.numpy()


def test_batched_uneven_meshes_match_individual_solves(gpu):
    rng = np.random.default_rng(0)
    systems = []
    for k, n in enumerate([5, 17, 9, 33, 12, 70]):
        mesh = wmesh.tri2(n, jitter=0.2, seed=k)
        f = rng.uniform(-1, 1, mesh['points'].shape[0])
        systems.append(wp.p1_system(mesh, None, f))
    bcg = BatchedCG(systems)
    R, timings = bcg.autotune(iters=3)
    assert R in timings
    infos = bcg.solve(rtol=1e-8)
    xs = bcg.solutions()
    for s, info, x in zip(systems, infos, xs):
        info_o, x_o = _oracle_solve(s, 1e-8)
        assert abs(info['iters'] - info_o['iters']) <= 1, (info, info_o)
        np.testing.assert_allclose(x, x_o, rtol=1e-6, atol=1e-9)
        assert info['residual'] <= 1e-8 * np.linalg.norm(s['b']) * 1.0000001
    # a second solve (graph replays) gives the same bits
    infos2 = bcg.solve(rtol=1e-8)
    assert infos2 == infos
    for a, b in zip(xs, bcg.solutions()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize('K', [8, 3, 64])
def test_batched_many_rhs_one_mesh(gpu, K):
    mesh = wmesh.tri2(24)
    systems = []
    for k in range(K):
        f = wp.rhs_uniform(mesh['points'].shape[0], k)
        systems.append(wp.p1_system(mesh, None, f))
    bcg = BatchedCG(systems)                       # shared operator
    assert bcg.multi is not None
    infos = bcg.solve(rtol=1e-8)
    xs = bcg.solutions()
    sep = BatchedCG(systems, share_operator=False)
    infos_sep = sep.solve(rtol=1e-8)
    xs_sep = sep.solutions()
    for k in (0, K // 2, K - 1):
        s, info = systems[k], infos[k]
        info_o, x_o = _oracle_solve(s, 1e-8)
        assert abs(info['iters'] - info_o['iters']) <= 1
        assert abs(info['iters'] - infos_sep[k]['iters']) <= 1
        assert info['residual'] <= 1e-8 * np.linalg.norm(s['b']) * 1.0000001
        np.testing.assert_allclose(xs[k], x_o, rtol=1e-6, atol=1e-9)
        np.testing.assert_allclose(xs[k], xs_sep[k], rtol=1e-6, atol=1e-9)
    # replaying the captured iterations again gives the same bits
    assert bcg.solve(rtol=1e-8) == infos
    for a, b in zip(xs, bcg.solutions()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize('jitter', [0.0, 0.2])
def test_multi_rhs_ell_product_gives_the_csr_bits(gpu, monkeypatch, jitter):
    """The shared-operator solver on the padded ELL copy of the operator
    (rows of at most 8 entries) and on the CSR: the same iterations,
    residuals and solution bits, for every right-hand side."""
    from easier_amd.numeric.solver.batched import MultiRhsCG
    mesh = wmesh.tri2(37, jitter=jitter, seed=3)
    systems = [wp.p1_system(mesh, None, wp.rhs_uniform(mesh['points'].shape[0], k))
               for k in range(13)]
    out = {}
    for ell in (True, False):
        monkeypatch.setattr(MultiRhsCG, 'use_ell', ell)
        bcg = BatchedCG(systems)
        assert bcg.multi is not None and (bcg.multi.ell is not None) == ell
        out[ell] = (bcg.solve(rtol=1e-10), bcg.solutions())
    assert out[True][0] == out[False][0]
    for a, b in zip(out[True][1], out[False][1]):
        assert np.array_equal(a, b)
    info_o, x_o = _oracle_solve(systems[5], 1e-10)
    assert abs(out[True][0][5]['iters'] - info_o['iters']) <= 1


def test_batched_split_autotuner_keeps_every_rhs(gpu):
    """The shared-operator autotuner (1, 2 or 4 groups of right-hand sides)
    picks a split by timing; whichever it picks, each right-hand side keeps
    its oracle iterations and solution."""
    mesh = wmesh.tri2(20)
    systems = [wp.p1_system(mesh, None, wp.rhs_uniform(mesh['points'].shape[0], k))
               for k in range(10)]
    bcg = BatchedCG(systems)
    S, timings = bcg.autotune(iters=3)
    assert S in (1, 2, 4) and set(timings) == {1, 2, 4}
    for S_forced in (4, S):
        bcg.groups = bcg._split(S_forced) if S_forced > 1 else [bcg.multi]
        infos, xs = bcg.solve(rtol=1e-8), bcg.solutions()
        assert len(infos) == len(xs) == 10
        for k in (0, 4, 9):
            info_o, x_o = _oracle_solve(systems[k], 1e-8)
            assert abs(infos[k]['iters'] - info_o['iters']) <= 1
            np.testing.assert_allclose(xs[k], x_o, rtol=1e-6, atol=1e-9)


@pytest.mark.parametrize('shared', [False, True])
def test_a_poisoned_instance_is_retired_by_its_mask(gpu, shared):
    """Fault injection: one instance's right-hand side holds a NaN. With
    maxiter=None its residual is NaN, so its own mask retires it (cg.py's
    NaN never converges; the batched kernels stop it instead of looping);
    solve() returns, and every healthy instance still matches its own
    oracle solve."""
    rng = np.random.default_rng(3)
    systems = []
    if shared:
        mesh = wmesh.tri2(16)
        systems = [wp.p1_system(mesh, None, wp.rhs_uniform(mesh['points'].shape[0], k))
                   for k in range(5)]
    else:
        for k, n in enumerate([6, 14, 9, 21, 11]):
            mesh = wmesh.tri2(n, jitter=0.2, seed=k)
            systems.append(wp.p1_system(mesh, None, rng.uniform(-1, 1, mesh['points'].shape[0])))
    bad = 2
    systems[bad] = dict(systems[bad], b=systems[bad]['b'].copy())
    systems[bad]['b'][len(systems[bad]['b']) // 2] = np.nan
    bcg = BatchedCG(systems)
    assert (bcg.multi is not None) == shared
    print(f'poisoned instance {bad}, shared={shared}', flush=True)
    infos = bcg.solve(rtol=1e-8, maxiter=None)
    xs = bcg.solutions()
    assert infos[bad]['residual'] != infos[bad]['residual']       # NaN
    for k, (s, info, x) in enumerate(zip(systems, infos, xs)):
        if k == bad:
            continue
        info_o, x_o = _oracle_solve(s, 1e-8)
        assert abs(info['iters'] - info_o['iters']) <= 1, (k, info, info_o)
        np.testing.assert_allclose(x, x_o, rtol=1e-6, atol=1e-9)
