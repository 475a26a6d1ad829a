"""The Python bridge binds every entry point with exactly the parameter list
the C header declares (a wrong arity is an error only at call time)."""
import ctypes
import os
import re

import pytest

from easier_amd import _native as nat
import easier_amd.numeric.solver.fused  # noqa: F401  (registers entry points)
import easier_amd.numeric.solver.mixed  # noqa: F401
import easier_amd.numeric.solver.batched  # noqa: F401
import easier_amd.numeric.solver.fcg  # noqa: F401
import easier_amd.numeric.preconditioner.learned  # noqa: F401
import easier_amd.numeric.preconditioner.multigrid  # noqa: F401

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declarations():
    decls = {}
    for fn in os.listdir(os.path.join(REPO, 'include')):
        if not fn.endswith('.h'):
            continue
        txt = open(os.path.join(REPO, 'include', fn)).read()
        txt = re.sub(r'/\*.*?\*/', ' ', txt, flags=re.S)
        for m in re.finditer(r'\b(?:int|const\s+char\s*\*)\s*(esr_\w+)\s*\(([^)]*)\)\s*;',
                             txt):
            params = m.group(2).strip()
            n = 0 if params in ('', 'void') else params.count(',') + 1
            decls[m.group(1)] = (n, params)
    return decls


def test_every_binding_matches_the_header():
    decls = _declarations()
    lib = nat.lib()
    sigs = dict(nat._SIGNATURES)
    sigs.update(nat._EXTRA_SIGNATURES)
    assert len(sigs) >= 40
    for name, (res, args) in sigs.items():
        assert name in decls, f'{name} is bound but not declared'
        n, params = decls[name]
        assert len(args) == n, f'{name}: bound {len(args)} args, header has {n}: {params}'
        assert getattr(lib, name).argtypes is not None


def test_every_declaration_is_bound():
    decls = _declarations()
    sigs = set(nat._SIGNATURES) | set(nat._EXTRA_SIGNATURES)
    unbound = sorted(set(decls) - sigs)
    assert not unbound, unbound
