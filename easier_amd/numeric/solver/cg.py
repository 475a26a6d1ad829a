# Portions of this file reproduce EASIER (jack-pan-ai/EASIER),
# Copyright (c) Microsoft Corporation, MIT License; the licence text is in
# easier_amd/THIRD_PARTY_NOTICES.txt.
"""Preconditioned conjugate gradients (easier/numeric/solver/cg.py:12-137).

Init/Step/Update are the reference's three esr.Modules with the same
forward bodies; solve() has the same control flow and counters:
  - tol = max(rtol * ||b||, atol) if atol else rtol * ||b||;
  - the loop stops on the recursive residual ||r|| <= tol (NaN never
    converges) or when iters reaches maxiter;
  - the converged step is not counted, so at most maxiter + 1 steps run.
After esr.compile(..., 'hip'), a CG whose operator is a Linsys and whose
preconditioner is None or preconditioner.Jacobi runs the fused device
iteration (solver/fused.py); anything else runs the compiled modules.
"""
from typing import Any, Dict, Optional

import torch
from torch import nn

from ...core import module as esr
from ...core.utils import logger


class Init(esr.Module):
    def __init__(self, A, b, x, M, r, z, p, bnorm) -> None:
        super().__init__()
        self.A = A
        self.x = x
        self.b = b
        self.M = M
        self.r = r
        self.z = z
        self.p = p
        self.bnorm = bnorm

    def forward(self):
        self.r[:] = self.b - self.A(self.x)
        self.z[:] = self.M(self.r)
        self.p[:] = self.z
        self.bnorm[:] = esr.norm(self.b).norm()


class Step(esr.Module):
    def __init__(self, A, x, r, z, p, rnorm, rzsum):
        super().__init__()
        self.A = A
        self.x = x
        self.p = p
        self.r = r
        self.z = z
        self.rnorm = rnorm
        self.rzsum = rzsum

    def forward(self):
        Ap = self.A(self.p)
        rzsum = esr.sum(self.r * self.z).sum()
        alpha = rzsum / esr.sum(self.p * Ap).sum()
        r1 = self.r - alpha * Ap
        self.x.add_(alpha * self.p)
        self.rzsum[:] = rzsum
        self.r[:] = r1
        self.rnorm[:] = esr.norm(r1).norm()


class Update(esr.Module):
    def __init__(self, x, M, r, z, p, rzsum):
        super().__init__()
        self.x = x
        self.M = M
        self.r = r
        self.z = z
        self.p = p
        self.rzsum = rzsum

    def forward(self):
        z1 = self.M(self.r)
        beta = esr.sum(self.r * z1).sum() / self.rzsum
        p1 = z1 + beta * self.p
        self.z[:] = z1
        self.p[:] = p1


def _identity(x):
    return x


def converged(rnorm: float, tol: float) -> bool:
    return rnorm == rnorm and rnorm <= tol


class CG(esr.Module):
    """Conjugate Gradient"""

    def __init__(self, A: nn.Module, b: esr.Tensor, x: esr.Tensor,
                 M: Optional[nn.Module] = None) -> None:
        super().__init__()
        self.A = A
        self.b = b
        self.x = x
        self.M = M if M else _identity
        self.r = esr.Tensor(esr.zeros_like(x), mode='partition')
        self.z = esr.Tensor(esr.zeros_like(x), mode='partition')
        self.p = esr.Tensor(esr.zeros_like(x), mode='partition')
        self.rzsum = esr.Tensor(torch.tensor([0.0], dtype=x.dtype),
                                mode='replicate')
        self.rnorm = esr.Tensor(torch.tensor([0.0], dtype=x.dtype),
                                mode='replicate')
        self.bnorm = esr.Tensor(torch.tensor([0.0], dtype=x.dtype),
                                mode='replicate')
        self.init = Init(A, b, x, self.M, self.r, self.z, self.p, self.bnorm)
        self.step = Step(A, x, self.r, self.z, self.p, self.rnorm, self.rzsum)
        self.update = Update(x, self.M, self.r, self.z, self.p, self.rzsum)
        self.easier_fused = None

    def easier_after_compile(self, prog):
        from .fused import make_fused_cg
        self.easier_fused = make_fused_cg(self, prog)

    def solve(self, rtol: float = 1e-5, atol: Optional[float] = None,
              maxiter: Optional[int] = None,
              debug_iter: Optional[int] = None) -> Dict[str, Any]:
        if self.easier_fused is not None:
            return self.easier_fused.solve(rtol, atol, maxiter, debug_iter)
        name = self.__class__.__name__
        self.init()
        rtol = rtol * float(self.bnorm.data.cpu()[0])
        tol = max(rtol, atol) if atol else rtol
        iters = 0
        while True:
            self.step()
            rnorm = float(self.rnorm.data.cpu()[0])
            if debug_iter is not None and iters % debug_iter == 0:
                logger.info(f"{name} residual {rnorm} at the {iters}-th iteration")
            if converged(rnorm, tol) or (maxiter is not None and iters >= maxiter):
                break
            iters += 1
            self.update()
        logger.info(f"{name} solver completed with residual {rnorm}"
                    f" at the {iters}-th iteration")
        return {'residual': rnorm, 'iters': iters}
