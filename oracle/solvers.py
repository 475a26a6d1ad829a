"""Linsys SpMV, CG and GMRES restated with torch CPU operators.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
"""
from typing import Callable, Optional

import numpy as np
import torch

from .ops import aggregate, reducer, selector


class Linsys:
    """numeric/linsys.py:23-24: reducer(selector(x) * Ae) + x * Av."""

    def __init__(self, Av, Ae, sel_idx, red_idx, n):
        self.Av, self.Ae = Av, Ae
        self.sel_idx, self.red_idx, self.n = sel_idx, red_idx, n

    def __call__(self, x):
        return reducer(selector(x, self.sel_idx) * self.Ae, self.red_idx,
                       self.n) + x * self.Av


def cg(A: Callable, b: torch.Tensor, x: torch.Tensor,
       M: Optional[Callable] = None, rtol=1e-5, atol=None, maxiter=None,
       final: Optional[dict] = None):
    """numeric/solver/cg.py:25-137. Returns (info, x) with info =
    {'residual', 'iters'}; x is updated in place. `final`, if given, receives
    the solver's r, p and rzsum when the loop exits (before the Update it
    skips)."""
    M = M if M else (lambda v: v)
    # Init (cg.py:25-29)
    r = b - A(x)
    z = M(r)
    p = z.clone()
    bnorm = aggregate('norm', b).norm()
    # solve (cg.py:112-137)
    rtol = rtol * bnorm
    tol = max(rtol, atol) if atol else rtol
    iters = 0
    while True:
        # Step (cg.py:44-52)
        Ap = A(p)
        rzsum = aggregate('sum', r * z).sum()
        alpha = rzsum / aggregate('sum', p * Ap).sum()
        r1 = r - alpha * Ap
        x.add_(alpha * p)
        r = r1
        rnorm = aggregate('norm', r1).norm()
        if (not torch.isnan(rnorm) and rnorm <= tol) or \
                (maxiter is not None and iters >= maxiter):
            break
        iters += 1
        # Update (cg.py:66-71)
        z1 = M(r)
        beta = aggregate('sum', r * z1).sum() / rzsum
        p = z1 + beta * p
        z = z1
    if final is not None:
        final.update(r=r, p=p, rzsum=float(rzsum))
    return {'residual': float(rnorm), 'iters': iters}, x


def gmres(A: Callable, b: torch.Tensor, x: torch.Tensor,
          M: Optional[Callable] = None, restart=20, rtol=1e-5, atol=None,
          maxiter=None):
    """numeric/solver/gmres.py:22-266 (H and B persist across cycles)."""
    M = M if M else (lambda v: v)
    n = x.shape[0]
    H = torch.zeros((restart + 1, restart), dtype=x.dtype)
    B = torch.zeros((restart + 1, 1), dtype=x.dtype)
    V = torch.zeros(tuple(x.shape) + (restart,), dtype=x.dtype)
    y = torch.zeros((restart, 1), dtype=x.dtype)
    bnorm = aggregate('norm', b).norm()
    rtol = rtol * bnorm
    tol = max(rtol, atol) if atol else rtol
    iters = 0
    while True:
        r = b - A(x)                                   # UpdateRnorm :22-25
        rnorm = aggregate('norm', r).norm()
        if (not torch.isnan(rnorm) and rnorm <= tol) or \
                (maxiter is not None and iters >= maxiter):
            break
        iters += 1
        B[0, 0] = rnorm
        V[:, ..., 0] = r / rnorm                       # InitV :48
        for j in range(restart):
            w = A(M(V[:, ..., j]))                     # InitW :64
            for i in range(j + 1):
                h = aggregate('sum', V[:, ..., i] * w).sum()   # SumW :79
                H[i, j] = h
                w = w - h * V[:, ..., i]               # UpdateW :105
            h = aggregate('norm', w).norm()            # NormW :90
            H[j + 1, j] = h
            if h < 1e-15:
                break
            elif j < restart - 1:
                V[:, ..., j + 1] = w / h               # UpdateV :120
        u, s, vt = torch.linalg.svd(H[:j + 2], full_matrices=False)
        y[:] = vt.transpose(0, 1) @ torch.diag_embed(
            1 / torch.clamp(s, min=1e-8)) @ u.transpose(0, 1) @ B[:j + 2]
        dx = torch.matmul(V[:, ..., :j + 1], y[:j + 1]).squeeze(dim=-1)
        x.add_(M(dx))                                  # UpdateX :133-136
    return {'residual': float(rnorm), 'iters': iters}, x


def dense_matrix(A: Linsys) -> np.ndarray:
    """The operator as a dense matrix (tiny systems only)."""
    n = A.n
    out = np.zeros((n, n))
    for k in range(n):
        e = torch.zeros(n, dtype=torch.float64)
        e[k] = 1.0
        out[:, k] = A(e).numpy()
    return out
