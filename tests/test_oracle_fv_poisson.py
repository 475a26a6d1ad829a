"""Pins of the finite-volume Poisson oracle (oracle/fv_poisson.py), the
operator of the reference's Poisson tutorial and integration test
(tutorial/poisson/assemble_poisson.py:160-230, tests/test_numeric.py:14-68):

* closed forms on the uniform 4-triangle mesh: neighbours across a
  half-diagonal have centroids d*sqrt(2)/3 apart and share a face of length
  d*sqrt(2)/2 (Af = 3/2); across a square side, d/3 and d (Af = 3); a wall
  face lies d/6 from its cell's centroid (wall term d / (d/6) = 6); so
  Ac = -6 inside and -9 at a wall;
* on a jittered mesh Af and Ac equal a brute-force construction from the two
  shared vertices of each cell pair (normal = rotated face vector pointing
  away from the source centroid, scaled by the face length);
* the operator is symmetric (Af of an edge equals Af of its reverse);
* on the uniform mesh it annihilates linear functions on interior cells
  (two-point fluxes are exact when the centroid line is normal to the face);
* b = rho * area with rho = exp(-200 |c - (0.5, 0.5)|^2)
  (assemble_poisson.py:215-219).
"""
import numpy as np
import pytest
import torch

from oracle import fv_poisson as ofv
from oracle import mesh as omesh
from oracle import ops as oops


def _jittered_tri4(n, amount, seed):
    m = omesh.tri4(n)
    pts = m['points'].copy()
    W = n + 1
    rng = np.random.default_rng(seed)
    interior = np.zeros(pts.shape[0], bool)
    for j in range(1, n):
        for i in range(1, n):
            interior[j * W + i] = True
    interior[W * W:] = True                      # the square centres
    pts[interior] += rng.uniform(-amount, amount, (int(interior.sum()), 2)) / n
    m['points'] = pts
    return m


def _brute(m):
    cells, pts = m['cells'], m['points']
    cent = pts[cells].mean(axis=1)
    nc = cells.shape[0]
    Af = np.zeros(m['src'].shape[0])
    for e, (s, d) in enumerate(zip(m['src'].tolist(), m['dst'].tolist())):
        a, b = sorted(set(cells[s].tolist()) & set(cells[d].tolist()))
        f = pts[b] - pts[a]
        nrm = np.array([f[1], -f[0]])
        if np.dot(nrm, 0.5 * (pts[a] + pts[b]) - cent[s]) < 0:
            nrm = -nrm
        dd = cent[d] - cent[s]
        Af[e] = np.dot(dd, nrm) / np.dot(dd, dd)
    Ac = np.zeros(nc)
    np.add.at(Ac, m['src'], -Af)
    for c, (a, b) in zip(m['bcells'].tolist(), m['bpoints'].tolist()):
        f = pts[b] - pts[a]
        nrm = np.array([f[1], -f[0]])
        mid = 0.5 * (pts[a] + pts[b])
        if np.dot(nrm, mid - cent[c]) < 0:
            nrm = -nrm
        dd = mid - cent[c]
        Ac[c] -= np.dot(dd, nrm) / np.dot(dd, dd)
    return Af, Ac


@pytest.mark.parametrize('n', [3, 10])
def test_uniform_mesh_closed_forms(n):
    m = omesh.tri4(n)
    a = ofv.assemble(m)
    Af, Ac = a['Af'].numpy(), a['Ac'].numpy()
    intra = np.arange(8 * n * n)                 # triangular_mesh.cpp:35-60
    np.testing.assert_allclose(Af[intra], 1.5, rtol=1e-13)
    np.testing.assert_allclose(Af[8 * n * n:], 3.0, rtol=1e-13)
    wall = np.zeros(Ac.shape[0], bool)
    wall[m['bcells']] = True
    np.testing.assert_allclose(Ac[~wall], -6.0, rtol=1e-13)
    np.testing.assert_allclose(Ac[wall], -9.0, rtol=1e-13)
    cent = m['points'][m['cells']].mean(axis=1)
    rho = np.exp(-200.0 * ((cent - 0.5) ** 2).sum(1))
    np.testing.assert_allclose(a['b'].numpy(), rho / (4 * n * n), rtol=1e-13)


def test_jittered_mesh_matches_brute_force():
    m = _jittered_tri4(7, 0.15, seed=11)
    a = ofv.assemble(m)
    Af, Ac = _brute(m)
    assert np.abs(Af - 1.5).max() > 0.05          # the jitter matters
    np.testing.assert_allclose(a['Af'].numpy(), Af, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(a['Ac'].numpy(), Ac, rtol=1e-12, atol=1e-12)


def test_operator_is_symmetric_and_annihilates_linear_functions():
    n = 8
    m = omesh.tri4(n)
    a = ofv.assemble(m)
    pos = {(s, d): e for e, (s, d) in enumerate(zip(m['src'].tolist(), m['dst'].tolist()))}
    Af = a['Af'].numpy()
    for (s, d), e in pos.items():
        assert Af[e] == Af[pos[(d, s)]]
    nc = m['cells'].shape[0]
    src, dst = torch.from_numpy(m['src']), torch.from_numpy(m['dst'])
    x = a['centroid'][:, 0] * 0.7 - a['centroid'][:, 1] * 1.3 + 0.2
    y = oops.reducer(oops.selector(x, dst) * a['Af'], src, nc) + x * a['Ac']
    wall = np.zeros(nc, bool)
    wall[m['bcells']] = True
    assert float(y[~torch.from_numpy(wall)].abs().max()) < 1e-12
    assert float(y[torch.from_numpy(wall)].abs().max()) > 0.1
