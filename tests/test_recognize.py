"""The fused CG/GMRES recognise their operator and preconditioner in the
traced graphs (numeric/solver/recognize.py), not by the class of the
objects: the reference's solvers take any operator and any callable M
(easier/numeric/solver/cg.py:74-104, gmres.py:139-194)."""
import pytest
import torch
from torch import nn

import easier_amd as esr
from easier_amd.core.compiler import compile_program
from easier_amd.numeric import Linsys
from easier_amd.numeric.preconditioner import Jacobi
from easier_amd.numeric.solver import CG, GMRES
from easier_amd.numeric.solver import recognize as rz
from easier_amd.numeric.solver.fused import FusedCG, FusedGMRES, make_fused_cg, \
    make_fused_gmres
from easier_amd.workloads import poisson as wp


class UserOperator(nn.Module):
    """The Linsys product written by a user, operands in another order."""

    def __init__(self, Av, Ae, sel, red):
        super().__init__()
        self.Av, self.Ae, self.sel, self.red = Av, Ae, sel, red

    def forward(self, x):
        return self.Av * x + self.red(self.Ae * self.sel(x))


class UserJacobi(nn.Module):
    def __init__(self, d):
        super().__init__()
        self.d = d

    def forward(self, r):
        return r / self.d


def _case(op_kind, m_kind, solver):
    _, sys_ = wp.config1_system(12, seed=0)
    eqn = wp.PoissonP1(sys_)
    A = eqn.A if op_kind == 'linsys' else UserOperator(eqn.Av, eqn.Ae, eqn.selector,
                                                        eqn.reducer)
    M = {'none': None, 'jacobi': Jacobi(eqn.Av), 'lambda': lambda r: r / eqn.Av,
         'module': UserJacobi(eqn.Av), 'scaled': lambda r: r * 0.5,
         'other_diag': lambda r: r / eqn.b}[m_kind]
    sol = CG(A, eqn.b, eqn.x, M) if solver == 'cg' else GMRES(A, eqn.b, eqn.x, M, restart=8)
    prog = compile_program([eqn, sol], 'rcm', torch.device('cpu'))
    return eqn, sol, prog


@pytest.mark.parametrize('solver', ['cg', 'gmres'])
@pytest.mark.parametrize('op_kind', ['linsys', 'user'])
@pytest.mark.parametrize('m_kind', ['none', 'jacobi', 'lambda', 'module'])
def test_fused_path_recognised_from_the_graph(solver, op_kind, m_kind):
    eqn, sol, prog = _case(op_kind, m_kind, solver)
    f = make_fused_cg(sol, prog) if solver == 'cg' else make_fused_gmres(sol, prog)
    assert isinstance(f, FusedCG if solver == 'cg' else FusedGMRES)
    assert f.op.Av.data_ptr() == eqn.Av.data.data_ptr()
    assert f.op.Ae.data_ptr() == eqn.Ae.data.data_ptr()
    if m_kind == 'none':
        assert f.diag is None
    else:
        assert f.diag.data_ptr() == eqn.Av.data.data_ptr()


@pytest.mark.parametrize('solver', ['cg', 'gmres'])
@pytest.mark.parametrize('m_kind', ['scaled', 'other_diag'])
def test_other_preconditioners_keep_the_compiled_modules(solver, m_kind):
    """r * 0.5 is not r / d; r / b divides by a tensor that is a partition
    vector of x's space too, so it is a valid Jacobi form and is fused."""
    eqn, sol, prog = _case('linsys', m_kind, solver)
    f = make_fused_cg(sol, prog) if solver == 'cg' else make_fused_gmres(sol, prog)
    if m_kind == 'scaled':
        assert f is None
    else:
        assert f is not None and f.diag.data_ptr() == eqn.b.data.data_ptr()


def test_match_linsys_rejects_a_different_product():
    """reducer(selector(x) * Ae) + x * Av only: a product with the diagonal
    term missing is not the operator."""
    _, sys_ = wp.config1_system(6, seed=0)
    eqn = wp.PoissonP1(sys_)

    class Partial(nn.Module):
        def __init__(self):
            super().__init__()
            self.sel, self.red, self.Ae = eqn.selector, eqn.reducer, eqn.Ae

        def forward(self, x):
            return self.red(self.sel(x) * self.Ae)
    sol = CG(Partial(), eqn.b, eqn.x)
    prog = compile_program([eqn, sol], 'rcm', torch.device('cpu'))
    assert rz.match_linsys(sol.step, rz.graph_of(prog, sol.step), sol.p) is None
    assert make_fused_cg(sol, prog) is None
