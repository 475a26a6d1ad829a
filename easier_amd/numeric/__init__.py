"""Numerics on top of esr.Modules (easier/numeric/__init__.py)."""
from . import solver
from . import preconditioner
from .linsys import Linsys
