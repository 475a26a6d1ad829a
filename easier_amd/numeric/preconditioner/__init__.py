"""Preconditioners for CG/GMRES (the reference's slot
easier/numeric/preconditioner/__init__.py:4 is empty)."""
from .jacobi import Jacobi
