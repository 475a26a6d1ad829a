"""End-to-end workloads through esr.compile(..., 'hip') against the oracle:

* the P1 Poisson system of the 100 x 100 benchmark configuration (2-triangle
  mesh, 10,201 vertices, f ~ U[-1, 1] from seed 0, zero Dirichlet),
  Jacobi-PCG, FP64, rtol 1e-8, maxiter 5000: iteration count equal to the
  oracle's (+-1 allowed by the solver contract), recursive residual within
  the tolerance, true residual recomputed on the host, and x within the
  distance two solutions stopped at those residuals can have
  (||A^-1|| (||r|| + ||r_o||), ||A^-1|| from shift-invert Lanczos);
* the reference's own integration workload (tests/test_numeric.py:14-68):
  the finite-volume Poisson tutorial (tutorial/poisson/assemble_poisson.py
  :80-230, poisson_main.py:16-53) on the 4-triangle n = 100 mesh (40,000
  cells), its operator assembled by an esr program on the GPU and compared
  with the oracle's assembly, then unpreconditioned CG and GMRES with
  atol = 1e-4, maxiter = 1000, compiling only the solver as the reference's
  test does; and once more to rtol 1e-10, where the comparison of iteration
  counts and solutions with the oracle is not trivial (solutions compared as
  above).
"""
import numpy as np
import pytest
import torch

import easier_amd as esr
from easier_amd.numeric.solver import CG, GMRES
from easier_amd.workloads import mesh as wmesh
from easier_amd.workloads import poisson as wp
from easier_amd.workloads import poisson_fv as wfv
from oracle import fv_poisson as ofv
from oracle import solvers as osol

pytestmark = pytest.mark.gpu


def _sparse(Av, Ae, red_idx, sel_idx):
    import scipy.sparse as sp
    n = len(Av)
    A = sp.csr_matrix((np.asarray(Ae), (np.asarray(red_idx), np.asarray(sel_idx))),
                      shape=(n, n))
    return (A + sp.diags(np.asarray(Av))).tocsc()


def _check_solution(A, b, x, x_o, label):
    """Both solutions stop at a residual, so they may differ by at most
    ||A^-1|| (||b - A x|| + ||b - A x_o||); ||A^-1|| = 1 / min |eigenvalue|
    of the symmetric operator (shift-invert Lanczos at 0)."""
    from scipy.sparse.linalg import eigsh
    b, x, x_o = (np.asarray(v, dtype=np.float64) for v in (b, x, x_o))
    res = float(np.linalg.norm(b - A @ x))
    res_o = float(np.linalg.norm(b - A @ x_o))
    lam = float(abs(eigsh(A, k=1, sigma=0, which='LM', return_eigenvectors=False)[0]))
    diff = float(np.linalg.norm(x - x_o))
    bound = 1.001 * (res + res_o) / lam + 1e-14 * float(np.linalg.norm(x_o))
    print(f'  {label}: |b - Ax| {res:.3e} (oracle {res_o:.3e}), |x - x_o| '
          f'{diff:.3e} <= {bound:.3e} (min |eig| {lam:.3e}), '
          f'relative {diff / float(np.linalg.norm(x_o)):.2e}', flush=True)
    assert diff <= bound
    return res


def test_config1_p1_poisson_jacobi_pcg(gpu):
    print('config 1: P1 100x100 Jacobi-PCG rtol 1e-8 maxiter 5000', flush=True)
    mesh, sys_ = wp.config1_system(100, seed=0)
    n = sys_['Av'].shape[0]
    assert n == 10201 and sys_['row'].shape[0] == 60400
    eqn = wp.PoissonP1(sys_)
    sol = CG(eqn.A, eqn.b, eqn.x, eqn.M)
    esr.compile([eqn, sol], 'hip')
    info = sol.solve(rtol=1e-8, maxiter=5000)
    x = eqn.x.collect()

    row, col = torch.from_numpy(sys_['row']), torch.from_numpy(sys_['col'])
    Av = torch.from_numpy(sys_['Av'])
    A_o = osol.Linsys(Av, torch.from_numpy(sys_['Ae']), col, row, n)
    b_o = torch.from_numpy(sys_['b'])
    info_o, x_o = osol.cg(A_o, b_o, torch.zeros(n, dtype=torch.float64),
                          M=lambda r: r / Av, rtol=1e-8, maxiter=5000)
    bnorm = float(b_o.norm())
    print(f'  iters {info["iters"]} (oracle {info_o["iters"]}), residual '
          f'{info["residual"]:.3e}', flush=True)
    assert abs(info['iters'] - info_o['iters']) <= 1
    assert info['residual'] <= 1e-8 * bnorm
    A = _sparse(sys_['Av'], sys_['Ae'], sys_['row'], sys_['col'])
    true_res = _check_solution(A, sys_['b'], x, x_o, 'config 1')
    assert true_res <= 2e-8 * bnorm


@pytest.fixture(scope='module')
def fv100():
    mesh = wmesh.tri4(100)
    assert mesh['cells'].shape[0] == 40000 and mesh['src'].shape[0] == 119600
    return mesh, ofv.assemble(mesh)


def test_fv_poisson_assembly_on_the_gpu_matches_the_oracle(gpu, fv100):
    print('FV Poisson tri4 n=100 assembly', flush=True)
    mesh, ref = fv100
    got = wfv.assemble(mesh)
    for k in ('Af', 'Ac', 'b', 'rho', 'area'):
        w = ref[k]
        err = float((got[k] - w).abs().max() / w.abs().max())
        assert err <= 1e-12, (k, err)
    assert float((got['cx'] - ref['centroid'][:, 0]).abs().max()) <= 1e-15


def _oracle_system(mesh, ref):
    nc = mesh['cells'].shape[0]
    A = osol.Linsys(ref['Ac'], ref['Af'], torch.from_numpy(mesh['dst']),
                    torch.from_numpy(mesh['src']), nc)
    return A, ref['b'], nc


@pytest.mark.parametrize('solver', ['cg', 'gmres'])
@pytest.mark.parametrize('tol', ['atol1e-4', 'rtol1e-10'])
def test_fv_poisson_integration_solve(gpu, fv100, solver, tol):
    print(f'FV Poisson tri4 n=100 {solver} {tol}', flush=True)
    mesh, ref = fv100
    eqn = wfv.FVPoisson(mesh, ref)
    sol = CG(eqn.A, eqn.b, eqn.x) if solver == 'cg' else GMRES(eqn.A, eqn.b, eqn.x)
    [sol] = esr.compile([sol], backend='hip')
    kw = dict(atol=1e-4, maxiter=1000) if tol == 'atol1e-4' else \
        dict(rtol=1e-10, maxiter=1000)
    info = sol.solve(**kw)
    x = eqn.x.collect()
    A_o, b_o, nc = _oracle_system(mesh, ref)
    run = osol.cg if solver == 'cg' else osol.gmres
    info_o, x_o = run(A_o, b_o, torch.zeros(nc, dtype=torch.float64), **kw)
    print(f'  iters {info["iters"]} (oracle {info_o["iters"]}), residual '
          f'{info["residual"]:.3e} (oracle {info_o["residual"]:.3e})', flush=True)
    if tol == 'atol1e-4':
        assert info['residual'] < 1e-4            # tests/test_numeric.py:25,40
    else:
        assert info['residual'] <= 1e-10 * float(b_o.norm())
    assert abs(info['iters'] - info_o['iters']) <= 1
    A = _sparse(ref['Ac'], ref['Af'], mesh['src'], mesh['dst'])
    _check_solution(A, b_o, x, x_o, f'{solver} {tol}')


class _UserJacobi(torch.nn.Module):
    def __init__(self, d):
        super().__init__()
        self.d = d

    def forward(self, r):
        return r / self.d


class _UserOperator(torch.nn.Module):
    """The Linsys product written by a user (operands in another order)."""

    def __init__(self, Av, Ae, sel, red):
        super().__init__()
        self.Av, self.Ae, self.sel, self.red = Av, Ae, sel, red

    def forward(self, x):
        return self.Av * x + self.red(self.Ae * self.sel(x))


@pytest.mark.parametrize('solver', ['cg', 'gmres'])
@pytest.mark.parametrize('kind', ['lambda', 'module', 'user_operator'])
def test_reference_style_callables_take_the_fused_path(gpu, kind, solver):
    """M = lambda r: r / A.Av, a user nn.Module M, or a user-written operator
    (the reference's solvers take any, cg.py:86, gmres.py:152): recognised
    in the traced graphs, solved by the fused device iteration, iterations
    and solution as the oracle's."""
    print(f'drop-in {kind} {solver}', flush=True)
    mesh, sys_ = wp.config1_system(40, seed=2)
    eqn = wp.PoissonP1(sys_)
    A = eqn.A if kind != 'user_operator' else _UserOperator(eqn.Av, eqn.Ae, eqn.selector,
                                                             eqn.reducer)
    M = _UserJacobi(eqn.Av) if kind == 'module' else (lambda r: r / eqn.Av)
    sol = CG(A, eqn.b, eqn.x, M) if solver == 'cg' else GMRES(A, eqn.b, eqn.x, M, restart=20)
    esr.compile([eqn, sol], 'hip')
    assert type(sol.easier_fused).__name__ == ('FusedCG' if solver == 'cg' else 'FusedGMRES')
    info = sol.solve(rtol=1e-9, maxiter=3000)
    n = sys_['Av'].shape[0]
    row, col = torch.from_numpy(sys_['row']), torch.from_numpy(sys_['col'])
    Av = torch.from_numpy(sys_['Av'])
    A_o = osol.Linsys(Av, torch.from_numpy(sys_['Ae']), col, row, n)
    b_o = torch.from_numpy(sys_['b'])
    run = osol.cg if solver == 'cg' else osol.gmres
    info_o, x_o = run(A_o, b_o, torch.zeros(n, dtype=torch.float64), M=lambda r: r / Av,
                      rtol=1e-9, maxiter=3000)
    print(f'  iters {info["iters"]} (oracle {info_o["iters"]})', flush=True)
    assert abs(info['iters'] - info_o['iters']) <= 1
    assert info['residual'] <= 1e-9 * float(b_o.norm())
    _check_solution(_sparse(sys_['Av'], sys_['Ae'], sys_['row'], sys_['col']), sys_['b'],
                    eqn.x.collect(), x_o, f'{kind} {solver}')
