"""Krylov solvers (easier/numeric/solver/__init__.py)."""
from .cg import CG
from .gmres import GMRES
