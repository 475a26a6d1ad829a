"""Krylov solvers (easier/numeric/solver/__init__.py), plus the MI355X
additions: mixed-precision refinement, batched and learned-preconditioned
flexible CG."""
from .cg import CG
from .gmres import GMRES
from .mixed import MixedPrecisionCG
from .batched import BatchedCG
from .fcg import FlexibleCG
