"""The user-facing documents stay in step with the code: the Python snippets
of INTEGRATION.md parse, and every C entry point the documents name is
declared in include/esr_hip.h (and bound by the Python bridge)."""
import ast
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DOCS = ('INTEGRATION.md', 'DESIGN.md', 'README.md')
# names that are not entry points: header/source file stems, prefixes and
# the generated fused kernels
NOT_ENTRIES = re.compile(r'^esr_(hip|krylov|host|common|fused_\w+|\w+_)$')


def _read(name):
    with open(os.path.join(REPO, name)) as f:
        return f.read()


def _declared():
    hdr = re.sub(r'/\*.*?\*/', ' ', _read(os.path.join('include', 'esr_hip.h')), flags=re.S)
    return set(re.findall(r'\b(esr_\w+)\s*\(', hdr))


def test_integration_snippets_parse():
    blocks = re.findall(r'```python\n(.*?)```', _read('INTEGRATION.md'), flags=re.S)
    assert blocks
    for b in blocks:
        ast.parse(b)


@pytest.mark.parametrize('doc', DOCS)
def test_named_entry_points_are_declared(doc):
    declared = _declared()
    named = {n for n in re.findall(r'\besr_[a-z0-9_]+\b', _read(doc))
             if not NOT_ENTRIES.match(n)}
    assert sorted(named - declared) == []


def test_declared_entry_points_are_bound():
    from easier_amd import _native as nat
    import easier_amd.numeric.solver.fused  # noqa: F401  (registers entry points)
    import easier_amd.numeric.solver.mixed  # noqa: F401
    import easier_amd.numeric.solver.batched  # noqa: F401
    import easier_amd.numeric.solver.fcg  # noqa: F401
    import easier_amd.numeric.preconditioner.learned  # noqa: F401
    import easier_amd.numeric.preconditioner.multigrid  # noqa: F401
    bound = set(nat._SIGNATURES) | set(nat._EXTRA_SIGNATURES)
    assert sorted(_declared() - bound) == []
