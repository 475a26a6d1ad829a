# Portions of this file reproduce EASIER (jack-pan-ai/EASIER),
# Copyright (c) Microsoft Corporation, MIT License; the licence text is in
# easier_amd/THIRD_PARTY_NOTICES.txt.
"""Restarted, right-preconditioned GMRES with modified Gram-Schmidt
(easier/numeric/solver/gmres.py:12-266).

Same esr.Modules and the same solve() control flow: a true residual at the
start of each restart cycle, `iters` counts cycles, an Arnoldi breakdown test
h < 1e-15, and the small least-squares problem solved through an SVD
pseudo-inverse with singular values clamped at 1e-8. The (j+2) x (j+1)
Hessenberg problem is solved on the host in float64 (it is 21 x 20 at most);
after esr.compile(..., 'hip') a GMRES on a Linsys runs its Arnoldi process
with fused device kernels (solver/fused.py).
"""
from typing import Any, Dict, Optional

import numpy as np
import torch
from torch import nn

from ...core import module as esr
from ...core import ops
from ...core.utils import logger


class UpdateRnorm(esr.Module):
    def __init__(self, A, b, x, r, rnorm):
        super().__init__()
        self.A = A
        self.b = b
        self.x = x
        self.r = r
        self.rnorm = rnorm

    def forward(self):
        r = self.b - self.A(self.x)
        self.r[:] = r
        self.rnorm[:] = esr.norm(r).norm()


class Init(esr.Module):
    def __init__(self, b, bnorm) -> None:
        super().__init__()
        self.b = b
        self.bnorm = bnorm

    def forward(self) -> None:
        self.bnorm[:] = esr.norm(self.b).norm()


class InitV(esr.Module):
    def __init__(self, r, rnorm, V):
        super().__init__()
        self.r = r
        self.rnorm = rnorm
        self.V = V

    def forward(self):
        self.V[:, ..., 0] = self.r / self.rnorm


class InitW(esr.Module):
    def __init__(self, A, M, V, w):
        super().__init__()
        self.A = A
        self.M = M
        self.V = V
        self.w = w
        self.j = esr.Tensor(torch.tensor([0], dtype=torch.int32),
                            mode='replicate')

    def forward(self):
        self.w[:] = self.A(self.M(self.V[:, ..., self.j].squeeze(-1)))


class SumW(esr.Module):
    def __init__(self, V, w, h):
        super().__init__()
        self.V = V
        self.w = w
        self.h = h
        self.i = esr.Tensor(torch.tensor([0], dtype=torch.int32),
                            mode='replicate')

    def forward(self):
        self.h[:] = esr.sum(self.V[:, ..., self.i].squeeze(-1) * self.w).sum()


class NormW(esr.Module):
    def __init__(self, w, h):
        super().__init__()
        self.w = w
        self.h = h

    def forward(self):
        self.h[:] = esr.norm(self.w).norm()


class UpdateW(esr.Module):
    def __init__(self, V, w, h):
        super().__init__()
        self.V = V
        self.w = w
        self.h = h
        self.i = esr.Tensor(torch.tensor([0], dtype=torch.int32),
                            mode='replicate')

    def forward(self):
        self.w.sub_(self.h * self.V[:, ..., self.i].squeeze(-1))


class UpdateV(esr.Module):
    def __init__(self, V, w, h):
        super().__init__()
        self.V = V
        self.w = w
        self.h = h
        self.i = esr.Tensor(torch.tensor([0], dtype=torch.int32),
                            mode='replicate')

    def forward(self):
        self.V[:, ..., self.i] = (self.w / self.h)[:, ..., None]


class UpdateX(esr.Module):
    def __init__(self, x, V, M, y, i):
        super().__init__()
        self.x = x
        self.V = V
        self.M = M
        self.y = y
        self.i = i

    def forward(self):
        V = self.V[:, ..., :self.i]
        dx = torch.matmul(V, self.y[:self.i]).squeeze(dim=-1)
        self.x.add_(self.M(dx))


def _identity(x):
    return x


def lstsq_svd(H: np.ndarray, B: np.ndarray) -> np.ndarray:
    """y = V diag(1/max(s, 1e-8)) U^T B for H = U diag(s) V^T, the thin SVD
    (gmres.py:255-258); H is H[:j+2] with all `restart` columns."""
    u, s, vt = np.linalg.svd(H, full_matrices=False)
    return vt.T @ np.diag(1.0 / np.maximum(s, 1e-8)) @ u.T @ B


class GMRES(esr.Module):
    """General Minimal Residual"""

    def __init__(self, A: nn.Module, b: esr.Tensor, x: esr.Tensor,
                 M: Optional[nn.Module] = None, restart=20):
        super().__init__()
        self.A = A
        self.b = b
        self.x = x
        self.M = M if M else _identity
        self.restart = restart
        self.r = esr.Tensor(esr.zeros_like(x), mode='partition')
        self.rnorm = esr.Tensor(torch.tensor([0.0], dtype=x.dtype),
                                mode='replicate')
        self.update_rnorm = UpdateRnorm(A, b, x, self.r, self.rnorm)
        self.bnorm = esr.Tensor(torch.tensor([0.0], dtype=x.dtype),
                                mode='replicate')
        self.init = Init(self.b, self.bnorm)
        self.H = esr.Tensor(torch.zeros((restart + 1, restart), dtype=x.dtype),
                            mode='replicate')
        self.B = esr.Tensor(torch.zeros((restart + 1, 1), dtype=x.dtype),
                            mode='replicate')
        self.V = esr.Tensor(esr.zeros(tuple(x.shape) + (restart,),
                                      dtype=x.dtype), mode='partition')
        self.init_V = InitV(self.r, self.rnorm, self.V)
        self.w = esr.Tensor(esr.zeros_like(x), mode='partition')
        self.init_w = InitW(A, self.M, self.V, self.w)
        self.h = esr.Tensor(torch.tensor([0.0], dtype=x.dtype),
                            mode='replicate')
        self.sum_w = SumW(self.V, self.w, self.h)
        self.update_w = UpdateW(self.V, self.w, self.h)
        self.norm_w = NormW(self.w, self.h)
        self.update_V = UpdateV(self.V, self.w, self.h)
        self.y = esr.Tensor(torch.zeros((restart, 1), dtype=x.dtype),
                            mode='replicate')
        self.update_x = nn.ModuleList(
            [UpdateX(self.x, self.V, self.M, self.y, i)
             for i in range(1, restart + 1)])
        # H and B persist across cycles and solve() calls, as the replicated
        # tensors of the reference do (only the entries a cycle writes
        # change; a cycle ended by breakdown keeps older columns).
        self.easier_host_H = np.zeros((restart + 1, restart))
        self.easier_host_B = np.zeros((restart + 1, 1))
        self.easier_fused = None

    def easier_after_compile(self, prog):
        from .fused import make_fused_gmres
        self.easier_fused = make_fused_gmres(self, prog)

    def _set_index(self, t: esr.Tensor, v: int):
        ops.fill_(t.data, v)

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
            self.update_rnorm()
            rnorm = float(self.rnorm.data.cpu()[0])
            if debug_iter is not None and iters % debug_iter == 0:
                logger.info(f"{name} residual {rnorm} at the {iters}-th iteration")
            if (rnorm == rnorm and rnorm <= tol) or \
                    (maxiter is not None and iters >= maxiter):
                break
            iters += 1
            H, B = self.easier_host_H, self.easier_host_B
            B[0, 0] = rnorm
            self.init_V()
            for j in range(self.restart):
                self._set_index(self.init_w.j, j)
                self.init_w()
                for i in range(j + 1):
                    self._set_index(self.sum_w.i, i)
                    self.sum_w()
                    H[i, j] = float(self.h.data.cpu()[0])
                    self._set_index(self.update_w.i, i)
                    self.update_w()
                self.norm_w()
                h = float(self.h.data.cpu()[0])
                H[j + 1, j] = h
                if h < 1e-15:
                    break
                elif j < self.restart - 1:
                    self._set_index(self.update_V.i, j + 1)
                    self.update_V()
            y = lstsq_svd(H[:j + 2], B[:j + 2])
            ops.copy_(self.y.data, torch.from_numpy(y).to(self.y.dtype)
                      .to(self.y.device))
            ops.copy_(self.H.data, torch.from_numpy(H).to(self.H.device))
            self.update_x[j]()
        logger.info(f"{name} solver completed with residual {rnorm}"
                    f" at the {iters}-th iteration")
        return {'residual': rnorm, 'iters': iters}
