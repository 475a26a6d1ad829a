"""Jacobi preconditioner M(r) = r / diag(A)."""
from torch import nn


class Jacobi(nn.Module):
    """M(r) = r / diag. Passed as `M` to CG/GMRES (cg.py:86, gmres.py:152 take
    any callable); as a plain nn.Module it is inlined into the traced
    forwards. For a Linsys A the diagonal is A.Av."""

    def __init__(self, diag) -> None:
        super().__init__()
        self.diag = diag

    def forward(self, r):
        return r / self.diag
