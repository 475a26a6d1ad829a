"""easier_amd: the EASIER mesh programming model on one AMD MI355X.

Import-compatible with the reference package `easier` (easier/__init__.py:4-14):

    import easier_amd as esr
    [eqn, sol] = esr.compile([eqn, sol], 'hip')
"""
from .core.module import (
    Tensor,
    Module,
    Selector, Reducer,
    sum, prod, norm, max, min,
    hdf5, full, zeros, ones, full_like, zeros_like, ones_like, arange,
)
from .core.jit import init, compile
from .core.cache import dump
from .core.utils import logger, EasierJitException
from . import numeric

__all__ = [
    'Tensor', 'Module', 'Selector', 'Reducer', 'sum', 'prod', 'norm', 'max',
    'min', 'hdf5', 'full', 'zeros', 'ones', 'full_like', 'zeros_like',
    'ones_like', 'arange', 'init', 'compile', 'dump', 'logger', 'EasierJitException',
    'numeric',
]
