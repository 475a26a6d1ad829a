"""The multigrid oracle pinned by what the mathematics fixes: P1
interpolation between nested meshes, the Galerkin identity of nested P1
spaces, scipy's transpose product, and the V-cycle's symmetry and
contraction."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import fem as ofem
from oracle import mesh as omesh
from oracle import multigrid as omg


def _locate(points, cells, x, y):
    """Index and barycentric weights of the triangle containing (x, y)."""
    for t, (a, b, c) in enumerate(cells):
        (xa, ya), (xb, yb), (xc, yc) = points[a], points[b], points[c]
        det = (xb - xa) * (yc - ya) - (xc - xa) * (yb - ya)
        l1 = ((xb - x) * (yc - y) - (xc - x) * (yb - y)) / det
        l2 = ((xc - x) * (ya - y) - (xa - x) * (yc - y)) / det
        l3 = 1.0 - l1 - l2
        if min(l1, l2, l3) >= -1e-12:
            return t, (l1, l2, l3)
    raise AssertionError('point outside the mesh')


@pytest.mark.parametrize('nf', [4, 8])
def test_prolongation_is_p1_interpolation(nf):
    nc = nf // 2
    cm, fm = omesh.tri2(nc), omesh.tri2(nf)
    rng = np.random.default_rng(nf)
    u = rng.standard_normal(cm['points'].shape[0])
    P = omg.prolongation(nf, keep_boundary=True)
    want = np.empty(fm['points'].shape[0])
    for i, (x, y) in enumerate(fm['points']):
        t, lam = _locate(cm['points'], cm['cells'], x, y)
        want[i] = sum(l * u[v] for l, v in zip(lam, cm['cells'][t]))
    np.testing.assert_allclose(P @ u, want, rtol=0, atol=1e-14)


def test_prolongation_drops_coarse_boundary():
    P = omg.prolongation(8)
    Pf = omg.prolongation(8, keep_boundary=True)
    inner = omg.interior(4)
    np.testing.assert_array_equal(P.toarray(), Pf.toarray() * inner[None, :])
    # fine boundary rows are empty (zero Dirichlet values)
    assert abs(P[~omg.interior(8)]).sum() == 0


@pytest.mark.parametrize('nf', [4, 8, 12])
def test_galerkin_equals_coarse_assembly(nf):
    """kappa constant on each coarse triangle: P^T A_h P = A_2h (nested P1
    spaces), boundary rows identity on both sides."""
    nc = nf // 2
    cm, fm = omesh.tri2(nc), omesh.tri2(nf)
    rng = np.random.default_rng(nf + 1)
    kc = np.exp(rng.standard_normal(cm['cells'].shape[0]))
    cen = fm['points'][fm['cells']].mean(axis=1)
    kf = np.array([kc[_locate(cm['points'], cm['cells'], x, y)[0]] for x, y in cen])
    fs = ofem.assemble(fm, kf)
    cs = ofem.assemble(cm, kc)
    A = sp.csr_matrix((fs['Ae'], (fs['row'], fs['col'])), shape=(fs['Av'].size,) * 2) + \
        sp.diags(fs['Av'])
    Ac_want = sp.csr_matrix((cs['Ae'], (cs['row'], cs['col'])), shape=(cs['Av'].size,) * 2) + \
        sp.diags(cs['Av'])
    Ac = omg.galerkin(A, omg.prolongation(nf), nc)
    np.testing.assert_allclose(Ac.toarray(), Ac_want.toarray(), rtol=0,
                               atol=1e-12 * abs(Ac_want).max())
    # the coarse operator lives on the coarse mesh's pattern
    rowptr, col, Ae, Av = omg.linsys_csr(Ac, nc)
    row = np.repeat(np.arange(rowptr.size - 1), np.diff(rowptr))
    rebuilt = sp.csr_matrix((Ae, (row, col)), shape=Ac.shape) + sp.diags(Av)
    np.testing.assert_allclose(rebuilt.toarray(), Ac.toarray(), rtol=0, atol=1e-15)


@pytest.mark.parametrize('nc', [1, 2, 5])
def test_pattern_is_the_p1_pattern(nc):
    s = ofem.assemble(omesh.tri2(nc))
    rowptr, col = omg.pattern(nc)
    want_ptr = np.concatenate([[0], np.cumsum(np.bincount(s['row'], minlength=(nc + 1) ** 2))])
    np.testing.assert_array_equal(rowptr, want_ptr)
    np.testing.assert_array_equal(col, s['col'])
    W = nc + 1
    assert rowptr[-1] == (W - 1) * (6 * W - 2)


def test_restrict_and_prolong_match_sparse_products():
    nf = 16
    rng = np.random.default_rng(3)
    fpos = rng.permutation((nf + 1) ** 2)
    fidx = np.empty_like(fpos)
    fidx[fpos] = np.arange(fpos.size)
    P = omg.prolongation(nf, fpos)
    t = rng.standard_normal(fpos.size)
    np.testing.assert_allclose(omg.restrict(t, nf, fidx), P.T @ t, rtol=1e-15, atol=1e-15)
    zc = rng.standard_normal((nf // 2 + 1) ** 2)
    np.testing.assert_allclose(omg.prolong_values(zc, nf, fpos), P @ zc, rtol=1e-15,
                               atol=1e-15)


def test_grid_positions_recover_jittered_grid():
    from easier_amd.workloads import mesh as wmesh
    n = 12
    m = wmesh.tri2(n, jitter=0.2, seed=5)
    pos = omg.grid_positions(m['points'], n)
    np.testing.assert_array_equal(pos, np.arange((n + 1) ** 2))
    perm = np.random.default_rng(0).permutation(pos.size)
    np.testing.assert_array_equal(omg.grid_positions(m['points'][perm], n), perm)


def _varcoef(n, seed):
    m = omesh.tri2(n)
    rng = np.random.default_rng(seed)
    s = ofem.assemble(m, np.exp(rng.standard_normal(m['cells'].shape[0])))
    A = sp.csr_matrix((s['Ae'], (s['row'], s['col'])), shape=(s['Av'].size,) * 2) + \
        sp.diags(s['Av'])
    return A, s


def test_vcycle_is_symmetric_and_contracts():
    n = 32
    A, s = _varcoef(n, 7)
    levels = omg.hierarchy(A, n, coarsest=4)
    assert [L.n for L in levels] == [32, 16, 8, 4]
    rng = np.random.default_rng(1)
    r1, r2 = rng.standard_normal((2, A.shape[0]))
    v1, v2 = omg.vcycle(levels, r1), omg.vcycle(levels, r2)
    assert abs(v1 @ r2 - r1 @ v2) < 1e-12 * abs(v1 @ r2)
    # stationary iteration x += V(b - A x): the error contracts by a factor
    # independent of n; damped Jacobi (w = 0.8) smooths high frequencies by
    # 0.6 per sweep, so a V(1,1) cycle gains about 0.6^2
    for kappa_const in (True, False):
        if kappa_const:
            ss = ofem.assemble(omesh.tri2(n))
            A = sp.csr_matrix((ss['Ae'], (ss['row'], ss['col'])), shape=A.shape) + \
                sp.diags(ss['Av'])
            levels = omg.hierarchy(A, n, coarsest=4)
        xs = rng.standard_normal(A.shape[0])
        b = A @ xs
        x = np.zeros_like(b)
        errs = []
        for _ in range(8):
            x = x + omg.vcycle(levels, b - A @ x)
            errs.append(np.linalg.norm(x - xs))
        rate = (errs[-1] / errs[2]) ** (1 / 5)
        assert rate < (0.45 if kappa_const else 0.6), rate


def test_two_level_with_exact_coarse_solve_kills_coarse_space():
    """A two-grid cycle with no smoothing is the A-orthogonal projection onto
    range(P): applying it to A P u returns P u exactly."""
    n = 8
    A, _ = _varcoef(n, 2)
    levels = omg.hierarchy(A, n, coarsest=4)
    assert len(levels) == 2
    u = np.random.default_rng(4).standard_normal(levels[1].A.shape[0])
    u[~omg.interior(4)] = 0
    L0, L1 = levels
    zc = L1.Ainv @ (L0.P.T @ (A @ (L0.P @ u)))
    np.testing.assert_allclose(L0.P @ zc, L0.P @ u, rtol=0, atol=1e-12 * abs(u).max())
