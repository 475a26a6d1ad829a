"""Host set-up of the algebraic multigrid levels (csrc/host_amg.cpp, smoothed
aggregation) against the oracle's restatement (oracle/amg.py: the same
aggregates, bit for bit) and scipy's sparse products: P = (I - omega D^-1 A)
P_tent with omega = 4 / (3 rho), R = P^T and the Galerkin operator P^T A P,
on variable-coefficient P1 systems of jittered meshes. The oracle's
aggregates are pinned by their defining properties. Also: the results do
not depend on the host thread count, and bad operators are refused."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg

from easier_amd.numeric.preconditioner.multigrid import amg_coarsen
from easier_amd.workloads import poisson as wp
from easier_amd.workloads import mesh as wmesh
from oracle import amg as oamg


def _system(n, seed=3, jitter=0.2):
    _, s, _ = wp.varcoef_system(n, seed, jitter=jitter)
    N = s['Av'].size
    rowptr = np.concatenate([[0], np.cumsum(np.bincount(s['row'], minlength=N))])
    return rowptr, s['col'], s['Ae'], s['Av']


def _matrix(rowptr, col, Ae, Av):
    n = Av.size
    return sp.csr_matrix((Ae, col, rowptr), shape=(n, n)) + sp.diags(Av)


def _strong(rowptr, col, Ae, Av, theta):
    n = Av.size
    row = np.repeat(np.arange(n), np.diff(rowptr))
    s = (Ae != 0) & (Ae * Ae >= theta ** 2 * np.abs(Av[row] * Av[col])) & (row != col)
    return sp.csr_matrix((np.ones(s.sum()), (row[s], col[s])), shape=(n, n))


@pytest.mark.parametrize('n,theta', [(37, 0.08), (53, 0.0), (24, 0.25)])
def test_level_matches_scipy(n, theta):
    ops = _system(n)
    A = _matrix(*ops)
    lv = amg_coarsen(*ops, theta=theta)
    m, nc = ops[3].size, lv['nc']
    assert 0 < nc < m // 3
    agg = lv['agg']
    S = _strong(*ops, theta)
    has_strong = np.diff(S.indptr) > 0
    # every vertex with a strong neighbour is in exactly one aggregate
    assert np.all((agg >= 0) == has_strong)
    assert np.array_equal(np.unique(agg[agg >= 0]), np.arange(nc))
    # rho: Gershgorin bound of D^-1 A
    rho = np.max((np.abs(A).sum(axis=1).A1) / ops[3])
    assert lv['rho'] == pytest.approx(rho, rel=1e-15)
    # P = (I - 4/(3 rho) D^-1 A) P_tent
    ok = agg >= 0
    T = sp.csr_matrix((np.ones(ok.sum()), (np.nonzero(ok)[0], agg[ok])), shape=(m, nc))
    P_ref = (sp.eye(m) - (4 / (3 * rho)) * sp.diags(1 / ops[3]) @ A) @ T
    P = sp.csr_matrix((lv['P'][2], lv['P'][1], lv['P'][0]), shape=(m, nc))
    assert abs(P - P_ref).max() <= 1e-14 * abs(P_ref).max()
    R = sp.csr_matrix((lv['R'][2], lv['R'][1], lv['R'][0]), shape=(nc, m))
    assert abs(R - P.T).max() == 0.0
    # columns ascending in every row of P, R and the coarse off-diagonal part
    for rp, cl in ((lv['P'][0], lv['P'][1]), (lv['R'][0], lv['R'][1]),
                   (lv['A'][0], lv['A'][1])):
        d = np.diff(cl.astype(np.int64))
        starts = np.zeros(cl.size, bool)
        starts[rp[:-1][np.diff(rp) > 0]] = True
        assert np.all((d > 0) | starts[1:])
    # coarse operator = P^T A P, SPD, diagonal apart
    crow, ccol, cAe, cAv = lv['A']
    assert np.all(ccol != np.repeat(np.arange(nc), np.diff(crow)))
    Ac = _matrix(crow, ccol, cAe, cAv)
    Ac_ref = (P.T @ A @ P).tocsr()
    assert abs(Ac - Ac_ref).max() <= 1e-13 * abs(Ac_ref).max()
    assert np.all(np.linalg.eigvalsh(Ac.toarray()) > 0)


def test_hierarchy_coarsens_to_a_few_rows():
    ops = _system(45)
    sizes = [ops[3].size]
    for lvl in range(10):
        lv = amg_coarsen(*ops, theta=0.08 * 0.5 ** lvl)
        if lv['nc'] == 0:
            break
        sizes.append(lv['nc'])
        ops = lv['A']
        if sizes[-1] < 64:
            break
    assert sizes[-1] < 64 and len(sizes) <= 5, sizes
    assert all(b < a / 3 for a, b in zip(sizes, sizes[1:])), sizes


def test_results_do_not_depend_on_the_thread_count():
    ops = _system(61)
    a = amg_coarsen(*ops, nthreads=1)
    b = amg_coarsen(*ops, nthreads=7)
    assert np.array_equal(a['agg'], b['agg'])
    for k in ('P', 'R', 'A'):
        for x, y in zip(a[k], b[k]):
            assert x.dtype == y.dtype and np.array_equal(x, y)


def test_isolated_rows_stay_out_of_the_aggregates():
    # Dirichlet rows of the P1 assembly are identity rows with no couplings
    _, s, _ = wp.varcoef_system(20, 1)
    N = s['Av'].size
    rowptr = np.concatenate([[0], np.cumsum(np.bincount(s['row'], minlength=N))])
    lv = amg_coarsen(rowptr, s['col'], s['Ae'], s['Av'])
    assert np.all(lv['agg'][s['boundary']] == -1)
    prow = lv['P'][0]
    assert np.all(np.diff(prow)[s['boundary']] == 0)


def test_refuses_bad_operators():
    rowptr, col, Ae, Av = _system(12)
    bad = Av.copy()
    bad[5] = 0.0
    with pytest.raises(ValueError, match='diagonal'):
        amg_coarsen(rowptr, col, Ae, bad)
    badc = col.copy()
    badc[3] = Av.size
    with pytest.raises(ValueError, match='column'):
        amg_coarsen(rowptr, badc, Ae, Av)
    with pytest.raises(ValueError, match='theta'):
        amg_coarsen(rowptr, col, Ae, Av, theta=1.5)
    # an indefinite operator: the coarse diagonal check fails
    with pytest.raises(ValueError, match='positive definite'):
        amg_coarsen(rowptr, col, -3.0 * np.abs(Ae), Av)
    with pytest.raises(ValueError, match='inconsistent'):
        amg_coarsen(rowptr[:-1], col, Ae, Av)


@pytest.mark.parametrize('n,theta', [(29, 0.08), (24, 0.25)])
def test_oracle_aggregates_have_their_defining_properties(n, theta):
    A = _matrix(*_system(n))
    S = oamg.strength(A, theta)
    agg, how = oamg.aggregate(A, theta)
    nb = [set(S.indices[S.indptr[i]:S.indptr[i + 1]]) for i in range(A.shape[0])]
    # aggregated exactly when the vertex has a strong neighbour
    assert np.array_equal(agg >= 0, np.array([bool(x) for x in nb]))
    nc = agg.max() + 1
    for a in range(nc):
        mem = np.nonzero(agg == a)[0]
        # connected in the strength graph
        sub = S[mem][:, mem]
        ncomp, _ = csg.connected_components(sub, directed=False)
        assert ncomp == 1
        p1 = [i for i in mem if how[i] == 1]
        if p1:
            # a pass-1 aggregate is a root plus exactly its strong neighbours
            roots = [i for i in p1 if set(p1) == {i} | nb[i]]
            assert roots, a
            # pass-2 members are strong neighbours of a pass-1 member
            for i in mem[how[mem] == 2]:
                assert nb[i] & set(p1)
        else:
            assert np.all(how[mem] == 3)
    # a vertex placed by pass 3 was free after pass 2: no strong neighbour of
    # it is in a pass-1 aggregate
    for i in np.nonzero(how == 3)[0]:
        assert not any(how[j] == 1 for j in nb[i])


@pytest.mark.parametrize('kind', ['tri2-jittered', 'tri4'])
def test_host_level_equals_oracle(kind):
    if kind == 'tri4':
        m = wmesh.tri4(14)
        rng = np.random.default_rng(2)
        s = wp.p1_system(m, np.exp(rng.standard_normal(m['cells'].shape[0])))
        N = s['Av'].size
        ops = (np.concatenate([[0], np.cumsum(np.bincount(s['row'], minlength=N))]),
               s['col'], s['Ae'], s['Av'])
    else:
        ops = _system(33, seed=8)
    A = _matrix(*ops)
    for lvl in range(3):
        theta = 0.08 * 0.5 ** lvl
        got = amg_coarsen(*ops, theta=theta)
        want = oamg.level(A, theta)
        assert np.array_equal(got['agg'], want['agg'])
        assert got['nc'] == want['nc']
        assert got['rho'] == pytest.approx(want['rho'], rel=1e-15)
        m, nc = A.shape[0], got['nc']
        P = sp.csr_matrix((got['P'][2], got['P'][1], got['P'][0]), shape=(m, nc))
        assert abs(P - want['P']).max() <= 1e-14 * abs(want['P']).max()
        Ac = _matrix(*got['A'])
        assert abs(Ac - want['Ac']).max() <= 1e-13 * abs(want['Ac']).max()
        ops, A = got['A'], Ac
        if nc < 40:
            break
