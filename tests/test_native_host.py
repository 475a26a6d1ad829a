"""libesrhip.so on the host: it loads, exports every entry point the C header
declares, and its host-side functions (mesh generators, stable CSR build,
RCM) agree with the oracle / library routines. No GPU needed."""
import os
import re

import numpy as np
import pytest

from easier_amd import _native as nat
from easier_amd.core import compiler
from easier_amd.workloads import mesh as wmesh
from oracle import mesh as omesh

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(REPO, 'include')):
        if fn.endswith('.h'):
            txt = open(os.path.join(REPO, 'include', fn)).read()
            names.update(re.findall(r'^\s*(?:const\s+)?\w+\s*\*?\s*(esr_\w+)\s*\(',
                                    txt, flags=re.M))
    return names


def test_library_exports_every_declared_symbol():
    lib = nat.lib()
    declared = _declared_symbols()
    assert len(declared) > 20
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    # and the Python bridge binds exactly what the header declares
    assert set(nat.exported_symbols()) <= declared
    assert lib.esr_abi_version() == 1


@pytest.mark.parametrize('n', [1, 2, 5, 9])
def test_native_tri4_equals_oracle(n):
    a, b = wmesh.tri4(n), omesh.tri4(n)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize('n', [1, 2, 5, 9])
def test_native_tri2_equals_oracle(n):
    a, b = wmesh.tri2(n), omesh.tri2(n)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_tri2_jitter_moves_only_interior_points():
    n = 8
    a, b = wmesh.tri2(n, jitter=0.2, seed=3), wmesh.tri2(n)
    d = np.abs(a['points'] - b['points'])
    W = n + 1
    I, J = np.meshgrid(np.arange(W), np.arange(W))
    interior = ((I > 0) & (I < n) & (J > 0) & (J < n)).ravel()
    assert np.all(d[~interior] == 0)
    assert np.all(d[interior] <= 0.2 / n + 1e-15) and d[interior].max() > 0
    assert np.array_equal(a['cells'], b['cells'])
    c = wmesh.tri2(n, jitter=0.2, seed=3)
    assert np.array_equal(a['points'], c['points'])     # seeded


def test_csr_from_idx_is_a_stable_sort():
    rng = np.random.default_rng(0)
    idx = rng.integers(0, 50, 1000)
    rowptr, perm = compiler.csr_from_idx(idx, 60)
    assert np.array_equal(perm, np.argsort(idx, kind='stable'))
    assert np.array_equal(np.diff(rowptr), np.bincount(idx, minlength=60))
    with pytest.raises(ValueError):
        compiler.csr_from_idx(np.array([0, 60]), 60)


def test_rcm_is_a_bandwidth_reducing_permutation():
    import scipy.sparse as sp
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    m = wmesh.tri2(12)
    rng = np.random.default_rng(1)
    nc = m['cells'].shape[0]
    shuffle = rng.permutation(nc)                 # scramble the cell numbering
    src, dst = shuffle[m['src']], shuffle[m['dst']]
    order = compiler.rcm_order(nc, src, dst)
    assert np.array_equal(np.sort(order), np.arange(nc))
    inv = np.empty_like(order)
    inv[order] = np.arange(nc)
    bw = compiler.bandwidth(inv[src], inv[dst])
    A = sp.csr_matrix((np.ones(src.shape[0]), (src, dst)), shape=(nc, nc))
    so = reverse_cuthill_mckee(A, symmetric_mode=True)
    sinv = np.empty_like(so)
    sinv[so] = np.arange(nc)
    bw_scipy = compiler.bandwidth(sinv[src], sinv[dst])
    assert bw <= bw_scipy + 2
    assert bw < compiler.bandwidth(src, dst) / 5


@pytest.mark.parametrize('mesh_kind,n,B', [('tri2', 40, 256), ('tri4', 23, 64), ('tri2', 7, 1000)])
def test_block_order_is_a_permutation_of_compact_blocks(mesh_kind, n, B):
    """esr_block_order: a permutation whose consecutive runs of B vertices
    are connected (each run is a breadth-first ball); on a mesh most edges
    stay inside a run, far more than in the RCM order."""
    m = getattr(wmesh, mesh_kind)(n)
    nc = m['cells'].shape[0]
    o = compiler.block_order(nc, m['src'], m['dst'], B)
    assert sorted(o.tolist()) == list(range(nc))
    pos = np.empty(nc, np.int64)
    pos[o] = np.arange(nc)
    ps, pd = pos[m['src']], pos[m['dst']]
    inside = np.mean(ps // B == pd // B)
    r = compiler.rcm_order(nc, m['src'], m['dst'])
    posr = np.empty(nc, np.int64)
    posr[r] = np.arange(nc)
    inside_rcm = np.mean(posr[m['src']] // B == posr[m['dst']] // B)
    if nc > B:
        assert inside > 0.7 and inside >= inside_rcm, (inside, inside_rcm)
    # runs are compact: mostly one connected piece (a run holds more only
    # where a ball ran out of unplaced vertices)
    adj = {}
    for a, b in zip(m['src'].tolist(), m['dst'].tolist()):
        adj.setdefault(a, []).append(b)
    pieces = []
    for k in range(0, nc, B):
        blk = set(o[k:k + B].tolist())
        left, count = set(blk), 0
        while left:
            count += 1
            stack = [left.pop()]
            while stack:
                u = stack.pop()
                for w in adj.get(u, []):
                    if w in left:
                        left.discard(w)
                        stack.append(w)
        pieces.append(count)
    assert np.mean(pieces) <= 2.0, pieces
