"""Pins the CPU oracle before anything is compared against it.

Each oracle function is checked against the reference's own known answers
(tests/core/test_module.py literals), the reference's own compiled mesh
generator where it builds (oracle/_ref), brute force on tiny inputs, library
routines (numpy.linalg, scipy.sparse) and invariants that a dropped term, a
wrong sign or a transposed index would break.
"""
import glob
import importlib.util
import math
import os

import numpy as np
import pytest
import torch

from oracle import mesh as omesh
from oracle import ops as oops
from oracle import solvers as osol
from oracle import swe as oswe

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---------------------------------------------------------------- Selector --
def test_selector_known_answer():
    # tests/core/test_module.py:29-39: selector(t) == t[idx]
    g = torch.Generator().manual_seed(0)
    idx = torch.randint(0, 10, (20,), generator=g)
    t = torch.randn((10, 1, 2), generator=g)
    out = oops.selector(t, idx)
    for e in range(20):
        assert torch.equal(out[e], t[int(idx[e])])


# ----------------------------------------------------------------- Reducer --
IDX = torch.tensor([1, 1, 2, 2, 2])


def test_reducer_known_answers():
    # tests/core/test_module.py:43-86
    edge = torch.randn((5, 3, 4), generator=torch.Generator().manual_seed(1))
    res0 = oops.reducer(edge, IDX, 3)
    res1 = torch.zeros((3, 3, 4))
    res1[1] = edge[0] + edge[1]
    res1[2] = edge[2] + edge[3] + edge[4]
    assert torch.equal(res0, res1)
    e1 = edge[:, 0, 0]
    r = oops.reducer(e1, IDX, 3)
    assert torch.equal(r, torch.tensor([0.0, e1[0] + e1[1], e1[2] + e1[3] + e1[4]]))
    out = torch.rand((3, 3, 4), generator=torch.Generator().manual_seed(2))
    keep = out.clone()
    res0 = oops.reducer(edge, IDX, 3, out=out)
    assert torch.equal(res0[0], keep[0])        # include_self=False keeps row 0
    assert torch.equal(res0[1], edge[0] + edge[1])
    assert torch.equal(res0[2], edge[2] + edge[3] + edge[4])


@pytest.mark.parametrize('reduce', ['sum', 'prod', 'mean', 'amax', 'amin'])
@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_reducer_matches_sequential_loop(reduce, dtype):
    g = torch.Generator().manual_seed(3)
    idx = torch.randint(0, 7, (40,), generator=g)
    t = torch.randn((40, 2), generator=g, dtype=torch.float64).to(dtype)
    if reduce in ('amax', 'amin'):
        t[5, 0] = float('nan')
    a = oops.reducer(t, idx, 9, reduce)
    b = oops.reducer_loop(t, idx, 9, reduce)
    assert torch.equal(torch.nan_to_num(a, nan=123.0), torch.nan_to_num(b, nan=123.0))


def test_reducer_sum_is_sequential_for_long_rows():
    # a single row of 3000 entries folds left to right from +0
    t = torch.randn(3000, dtype=torch.float64, generator=torch.Generator().manual_seed(4))
    idx = torch.zeros(3000, dtype=torch.long)
    acc = 0.0
    for v in t.tolist():
        acc += v
    assert oops.reducer(t, idx, 1).item() == acc


# ------------------------------------------------------------- aggregators --
def test_aggregators_against_fsum():
    t = torch.randn((5000, 3), dtype=torch.float64,
                    generator=torch.Generator().manual_seed(5))
    s = oops.aggregate('sum', t)
    assert s.shape == (1, 3)
    for c in range(3):
        exact = math.fsum(t[:, c].tolist())
        bound = 5000 * 2.2e-16 * float(t[:, c].abs().sum())
        assert abs(float(s[0, c]) - exact) <= bound
    nrm = oops.aggregate('norm', t)
    for c in range(3):
        assert abs(float(nrm[0, c]) - math.sqrt(math.fsum(v * v for v in t[:, c].tolist()))) \
            <= 1e-13 * float(nrm[0, c])
    assert torch.equal(oops.aggregate('max', t)[0], t.max(0).values)
    assert torch.equal(oops.aggregate('min', t)[0], t.min(0).values)


# -------------------------------------------------------------------- mesh --
def _load_ref():
    so = glob.glob(os.path.join(REPO, 'oracle', '_ref', 'easier_ref_C*.so'))
    if not so:
        return None
    spec = importlib.util.spec_from_file_location('easier_ref_C', so[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.mark.parametrize('n', [1, 2, 3, 5, 16])
def test_tri4_matches_reference_build(n):
    ref = _load_ref()
    if ref is None:
        pytest.skip('oracle/_ref not built (make -C oracle ref)')
    torch.set_num_threads(1)  # the reference's OpenMP loop races on shared points
    src, dst, cells, bcells, bpoints, points = ref.generate_triangular_mesh(n)
    m = omesh.tri4(n)
    assert np.array_equal(m['src'], src.numpy())
    assert np.array_equal(m['dst'], dst.numpy())
    assert np.array_equal(m['cells'].ravel(), cells.numpy())
    assert np.array_equal(m['bcells'], bcells.numpy())
    assert np.array_equal(m['bpoints'].ravel(), bpoints.numpy())
    assert np.array_equal(m['points'].ravel(), points.numpy())


@pytest.mark.parametrize('gen,n', [(omesh.tri4, 3), (omesh.tri4, 6),
                                   (omesh.tri2, 3), (omesh.tri2, 7)])
def test_mesh_structure_brute_force(gen, n):
    m = gen(n)
    k = 4 if gen is omesh.tri4 else 2
    assert m['cells'].shape[0] == k * n * n
    # triangular_mesh.cpp:14-16 (tri4); 6n^2-4n directed dual edges (tri2)
    assert m['src'].shape[0] == (12 * n * n - 4 * n if k == 4 else 6 * n * n - 4 * n)
    pairs, boundary = omesh.dual_edges_from_cells(m['cells'])
    got = set(zip(m['src'].tolist(), m['dst'].tolist()))
    assert got == pairs and len(got) == m['src'].shape[0]
    bf = {(int(c), (min(a, b), max(a, b))) for c, (a, b) in
          zip(m['bcells'], m['bpoints'].tolist())}
    assert bf == boundary
    p = m['points']
    c = m['cells']
    area = 0.5 * ((p[c[:, 1], 0] - p[c[:, 0], 0]) * (p[c[:, 2], 1] - p[c[:, 0], 1])
                  - (p[c[:, 2], 0] - p[c[:, 0], 0]) * (p[c[:, 1], 1] - p[c[:, 0], 1]))
    assert np.all(area > 0)                       # counter-clockwise
    assert abs(area.sum() - 1.0) < 1e-12          # tiles the unit square


# ------------------------------------------------------------------ solvers --
def _grid_laplacian(m):
    """5-point Laplacian on an m x m interior grid as a Linsys."""
    n = m * m
    rows, cols = [], []
    for j in range(m):
        for i in range(m):
            v = j * m + i
            for di, dj in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                a, b = i + di, j + dj
                if 0 <= a < m and 0 <= b < m:
                    rows.append(v)
                    cols.append(b * m + a)
    Ae = -torch.ones(len(rows), dtype=torch.float64)
    Av = torch.full((n,), 4.0, dtype=torch.float64)
    return osol.Linsys(Av, Ae, torch.tensor(cols), torch.tensor(rows), n)


def test_linsys_matches_scipy_csr():
    import scipy.sparse as sp
    A = _grid_laplacian(5)
    x = torch.randn(25, dtype=torch.float64, generator=torch.Generator().manual_seed(6))
    M = sp.csr_matrix((A.Ae.numpy(), (A.red_idx.numpy(), A.sel_idx.numpy())),
                      shape=(25, 25)) + sp.diags(A.Av.numpy())
    np.testing.assert_allclose(A(x).numpy(), M @ x.numpy(), rtol=1e-15, atol=1e-15)
    # closed-form eigenpair of the Dirichlet Laplacian: sin(j pi x) sin(k pi y)
    m, h = 5, 1.0 / 6
    jj, kk = 2, 3
    g = np.array([[math.sin(jj * math.pi * (i + 1) * h) * math.sin(kk * math.pi * (j + 1) * h)
                   for i in range(m)] for j in range(m)]).ravel()
    lam = 4 - 2 * math.cos(jj * math.pi * h) - 2 * math.cos(kk * math.pi * h)
    np.testing.assert_allclose(A(torch.from_numpy(g)).numpy(), lam * g, atol=1e-13)


def test_cg_exact_on_tiny_system():
    A = _grid_laplacian(3)
    b = torch.randn(9, dtype=torch.float64, generator=torch.Generator().manual_seed(7))
    x = torch.zeros(9, dtype=torch.float64)
    info, x = osol.cg(A, b, x, rtol=1e-12, maxiter=100)
    xs = np.linalg.solve(osol.dense_matrix(A), b.numpy())
    np.testing.assert_allclose(x.numpy(), xs, rtol=1e-10, atol=1e-12)
    assert info['iters'] <= 9                      # <= n steps in exact arithmetic
    # preconditioned by the diagonal: same solution
    x2 = torch.zeros(9, dtype=torch.float64)
    info2, x2 = osol.cg(A, b, x2, M=lambda r: r / A.Av, rtol=1e-12, maxiter=100)
    np.testing.assert_allclose(x2.numpy(), xs, rtol=1e-10, atol=1e-12)


def test_cg_counts_like_the_reference():
    # maxiter caps the counted iterations; the converged step is not counted
    A = _grid_laplacian(6)
    b = torch.ones(36, dtype=torch.float64)
    info, _ = osol.cg(A, b, torch.zeros(36, dtype=torch.float64), rtol=1e-30, maxiter=3)
    assert info['iters'] == 3
    info, _ = osol.cg(A, b, torch.zeros(36, dtype=torch.float64), rtol=1e-10)
    info1, _ = osol.cg(A, b, torch.zeros(36, dtype=torch.float64), rtol=1e-10,
                       maxiter=info['iters'])
    assert info1['iters'] == info['iters'] and info1['residual'] == info['residual']


def test_gmres_one_cycle_when_restart_covers_n():
    A = _grid_laplacian(3)
    b = torch.randn(9, dtype=torch.float64, generator=torch.Generator().manual_seed(8))
    x = torch.zeros(9, dtype=torch.float64)
    info, x = osol.gmres(A, b, x, restart=20, rtol=1e-10, maxiter=50)
    xs = np.linalg.solve(osol.dense_matrix(A), b.numpy())
    np.testing.assert_allclose(x.numpy(), xs, rtol=1e-8, atol=1e-10)
    assert info['iters'] == 1
    x = torch.zeros(9, dtype=torch.float64)
    info, x = osol.gmres(A, b, x, restart=3, atol=1e-4, maxiter=1000)
    assert info['residual'] < 1e-4                 # tests/test_numeric.py:40


# ---------------------------------------------------------------------- SWE --
def test_swe_geometry_invariants():
    m = omesh.tri4(4)
    geo = oswe.geometry(m)
    src, dst = m['src'], m['dst']
    pos = {(a, b): e for e, (a, b) in enumerate(zip(src.tolist(), dst.tolist()))}
    for (a, b), e in pos.items():
        r = pos[(b, a)]
        assert geo['sx'][e] == -geo['sx'][r] and geo['sy'][e] == -geo['sy'][r]
        assert 0 < geo['alpha'][e] < 1
    # closed cells: interior outward normals plus wall normals sum to zero
    nc = m['cells'].shape[0]
    sx = oops.reducer(geo['sx'], torch.from_numpy(src), nc) - \
        oops.reducer(geo['bsx'], torch.from_numpy(m['bcells']), nc)
    sy = oops.reducer(geo['sy'], torch.from_numpy(src), nc) - \
        oops.reducer(geo['bsy'], torch.from_numpy(m['bcells']), nc)
    assert float(sx.abs().max()) < 1e-15 and float(sy.abs().max()) < 1e-15
    assert abs(float(geo['area'].sum()) - 1.0) < 1e-13


@pytest.mark.parametrize('gen', [omesh.tri4, omesh.tri2])
def test_swe_lake_at_rest_and_mass(gen):
    m = gen(6)
    geo = oswe.geometry(m)
    geo_flat = dict(geo)
    geo_flat['h'] = torch.full_like(geo['h'], 1.5)
    s = oswe.SWE(m, geo_flat, dt=0.01, g=9.81)
    for _ in range(3):
        s()
    assert float((s.h - 1.5).abs().max()) < 1e-13
    assert float(s.uh.abs().max()) < 1e-13 and float(s.vh.abs().max()) < 1e-13
    s = oswe.SWE(m, oswe.geometry(m, centre=(0.5, 0.5)), dt=0.002, g=1.0)
    mass0 = float((s.h * s.area).sum())
    for _ in range(10):
        s()
    assert abs(float((s.h * s.area).sum()) - mass0) < 1e-14
    assert float(s.uh.abs().max()) > 0
