"""P1 assembly: the oracle pinned by closed forms, then the native assembler
against the oracle; the Gaussian random field against its covariance."""
import math

import numpy as np
import pytest

from easier_amd.workloads import mesh as wmesh
from easier_amd.workloads import poisson as wp
from oracle import fem as ofem
from oracle import mesh as omesh


def test_uniform_mesh_is_the_five_point_stencil():
    n = 6
    m = omesh.tri2(n)
    s = ofem.assemble(m, None, np.ones(m['points'].shape[0]))
    W = n + 1
    h = 1.0 / n
    for v in range(W * W):
        I, J = v % W, v // W
        if s['boundary'][v]:
            assert 0 in (I, J) or n in (I, J)
            assert s['Av'][v] == 1.0 and s['b'][v] == 0.0
            continue
        assert abs(s['Av'][v] - 4.0) < 1e-14
        cols = s['col'][s['row'] == v]
        vals = s['Ae'][s['row'] == v]
        for c, a in zip(cols, vals):
            ci, cj = c % W, c // W
            nb_boundary = s['boundary'][c]
            if abs(ci - I) + abs(cj - J) == 1 and not nb_boundary:
                assert abs(a + 1.0) < 1e-14          # axis neighbour: -1
            else:
                assert a == 0.0                        # diagonal: cot 90 = 0
        assert abs(s['b'][v] - h * h) < 1e-15       # mass load of f = 1


def test_constant_in_kernel_and_symmetry():
    m = omesh.tri2(5)
    m['points'] = wmesh.tri2(5, jitter=0.2, seed=4)['points']
    rng = np.random.default_rng(0)
    kap = np.exp(rng.standard_normal(m['cells'].shape[0]))
    s = ofem.assemble(m, kap, None)
    # the operator before Dirichlet elimination annihilates constants: check on
    # rows whose neighbours are all interior (their row sum is untouched)
    for v in np.nonzero(~s['boundary'])[0]:
        sel = s['row'] == v
        if not s['boundary'][s['col'][sel]].any():
            assert abs(s['Av'][v] + s['Ae'][sel].sum()) < 1e-12 * s['Av'][v]
    key = {(r, c): a for r, c, a in zip(s['row'], s['col'], s['Ae'])}
    for (r, c), a in key.items():
        assert key[(c, r)] == a


@pytest.mark.parametrize('n,jitter', [(4, 0.0), (7, 0.2)])
def test_native_assembler_matches_oracle(n, jitter):
    m = wmesh.tri2(n, jitter=jitter, seed=1)
    rng = np.random.default_rng(n)
    kap = np.exp(rng.standard_normal(m['cells'].shape[0]))
    f = rng.uniform(-1, 1, m['points'].shape[0])
    a = wp.p1_system(m, kap, f)
    b = ofem.assemble(m, kap, f)
    for k in ('row', 'col', 'boundary'):
        assert np.array_equal(a[k], b[k]), k
    for k in ('Ae', 'Av', 'b'):
        assert np.array_equal(a[k], b[k]), k      # same operation order


def test_config1_sizes():
    mesh, sys_ = wp.config1_system(100, seed=0)
    assert mesh['points'].shape[0] == 10201
    assert mesh['cells'].shape[0] == 20000
    assert sys_['row'].shape[0] == 2 * (3 * 100 * 100 + 2 * 100)   # 60,400


def test_gaussian_random_field_statistics():
    n, ell = 64, 0.05
    zs = np.stack([wp.gaussian_random_field(n, s, ell=ell) for s in range(40)])
    assert zs.shape == (40, n + 1, n + 1)
    var = zs.var()
    assert abs(var - 1.0) < 0.1
    lag = int(round(ell * n))                    # one correlation length
    c = (zs[:, :, :-lag] * zs[:, :, lag:]).mean()
    assert abs(c - math.exp(-0.5 * (lag / n / ell) ** 2)) < 0.1
    assert np.array_equal(wp.gaussian_random_field(n, 3), wp.gaussian_random_field(n, 3))


def test_load_vectors_equal_the_assemblers_bits():
    """workloads.poisson.load_vectors (many right-hand sides of one mesh in
    one pass) gives the load vector esr_p1_assemble computes, bit for bit, on
    a structured and a jittered mesh, boundary rows zero."""
    import numpy as np
    from easier_amd.workloads import mesh as wmesh
    from easier_amd.workloads import poisson as wp
    for m in (wmesh.tri2(9), wmesh.tri2(33, jitter=0.2, seed=5)):
        nv = m['points'].shape[0]
        fs = [wp.rhs_uniform(nv, k) for k in range(4)]
        base = wp.p1_system(m, None, None)
        B = wp.load_vectors(m, fs, base['boundary'])
        for k in range(4):
            want = wp.p1_system(m, None, fs[k])['b']
            assert np.array_equal(B[k], want)
        assert not B[:, base['boundary']].any()
