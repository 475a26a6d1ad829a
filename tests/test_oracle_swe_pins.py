"""Independent pins of the shallow-water oracle (oracle/swe.py).

The oracle restates tutorial/shallow_water_equation/assemble_shallow_water.py
:154-253 (geometry) and swe_main.py:86-141 (the RK4 step). The checks here do
not re-evaluate those formulas; they compute the same quantities another way
or use properties the scheme must have:

* geometry: alpha is 1/2 on every face of the uniform 2- and 4-triangle
  meshes (each pair of neighbours is the mirror or point image of the other
  across the shared face); on a jittered mesh alpha, the length-scaled normals
  sx, sy and the wall normals bsx, bsy equal a brute-force construction from
  the shared vertex pair of each cell pair (segment/line intersection solved
  with numpy.linalg.solve, normal = the rotated face vector pointing away from
  the source centroid); areas equal |cross| / 2.
* delta (the right-hand side): a uniform state (h, uh, vh) = (H, U, V) has the
  closed form delta_h = -(U Wx + V Wy)/A, delta_uh = -(U^2/H Wx + U V/H Wy)/A,
  delta_vh = -(U V/H Wx + V^2/H Wy)/A per cell, with W the sum of the cell's
  outward wall normals (length-scaled, brute force from the wall points):
  the reference scatters into dst with the src-outward normal, so the
  interior normals of a cell sum to +W, and its wall term adds -W * H^2/2.
* the full step is equivariant under the symmetries of the 4-triangle mesh
  (90-degree rotation and reflection of the unit square): a run started from
  a rotated Gaussian equals the rotated run, cell for cell, with the momentum
  vector rotated. An x/y transposition of the normals or of u_f/v_f breaks
  this (the planted variants below must fail the same check).
* the RK4 combine: with delta replaced by the linear map U -> aU one step
  multiplies U by 1 + 2z/3 + z^2/3 + z^3/8 + z^4/24 (z = a dt), the
  polynomial of the reference's stage inputs (U + dt/2 d1, U + dt/2 d2,
  U + dt d3) and combine dt/6 (d1 + d2 + d3 + d4) (swe_main.py:121-141);
  classical RK4 weights (1, 2, 2, 1) give 1 + z + z^2/2 + z^3/6 + z^4/24 and
  fail.
"""
import numpy as np
import pytest
import torch

from oracle import mesh as omesh
from oracle import swe as oswe


def _jittered_tri2(n, amount, seed):
    m = omesh.tri2(n)
    pts = m['points'].copy()
    W = n + 1
    rng = np.random.default_rng(seed)
    for j in range(1, n):
        for i in range(1, n):
            pts[j * W + i] += rng.uniform(-amount, amount, 2) / n
    m['points'] = pts
    return m


def _brute_geometry(m):
    """Per directed edge: alpha, sx, sy; per wall face: bsx, bsy; per cell:
    area -- all from the shared vertices of each cell pair."""
    cells, pts = m['cells'], m['points']
    cent = pts[cells].mean(axis=1)
    ne = m['src'].shape[0]
    alpha, sx, sy = np.zeros(ne), np.zeros(ne), np.zeros(ne)
    for e, (s, d) in enumerate(zip(m['src'].tolist(), m['dst'].tolist())):
        shared = sorted(set(cells[s].tolist()) & set(cells[d].tolist()))
        assert len(shared) == 2, (e, s, d)
        p0, p1 = pts[shared[0]], pts[shared[1]]
        # c_s + t (c_d - c_s) = p0 + u (p1 - p0)
        A = np.array([cent[d] - cent[s], p0 - p1]).T
        t, _ = np.linalg.solve(A, p0 - cent[s])
        alpha[e] = t
        f = p1 - p0
        nrm = np.array([f[1], -f[0]])
        if np.dot(nrm, 0.5 * (p0 + p1) - cent[s]) < 0:
            nrm = -nrm
        sx[e], sy[e] = nrm
    nb = m['bcells'].shape[0]
    bsx, bsy = np.zeros(nb), np.zeros(nb)
    for k, (c, (a, b)) in enumerate(zip(m['bcells'].tolist(), m['bpoints'].tolist())):
        f = pts[b] - pts[a]
        nrm = np.array([f[1], -f[0]])
        if np.dot(nrm, 0.5 * (pts[a] + pts[b]) - cent[c]) < 0:
            nrm = -nrm
        bsx[k], bsy[k] = -nrm                  # the reference stores -outward
    e1 = pts[cells[:, 1]] - pts[cells[:, 0]]
    e2 = pts[cells[:, 2]] - pts[cells[:, 0]]
    area = 0.5 * np.abs(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])
    return dict(alpha=alpha, sx=sx, sy=sy, bsx=bsx, bsy=bsy, area=area)


# ------------------------------------------------------------------ geometry --
@pytest.mark.parametrize('gen,n', [(omesh.tri2, 5), (omesh.tri2, 16),
                                   (omesh.tri4, 5), (omesh.tri4, 12)])
def test_alpha_is_one_half_on_uniform_meshes(gen, n):
    geo = oswe.geometry(gen(n))
    a = geo['alpha'].numpy()
    # the formula cancels coordinates of size 1 down to differences of size
    # 1/n, so its rounding error grows like n * eps
    assert np.abs(a - 0.5).max() <= 16 * n * np.finfo(np.float64).eps


@pytest.mark.parametrize('mesh', ['tri2_jitter', 'tri4'])
def test_geometry_matches_brute_force(mesh):
    m = _jittered_tri2(9, 0.2, seed=3) if mesh == 'tri2_jitter' else omesh.tri4(6)
    geo = oswe.geometry(m)
    ref = _brute_geometry(m)
    for k in ('alpha', 'sx', 'sy', 'bsx', 'bsy', 'area'):
        got = geo[k].numpy()
        np.testing.assert_allclose(got, ref[k], rtol=1e-12, atol=1e-14, err_msg=k)
    if mesh == 'tri2_jitter':                       # alpha is not trivially 1/2
        assert np.abs(ref['alpha'] - 0.5).max() > 0.02


# --------------------------------------------------------------------- delta --
@pytest.mark.parametrize('gen', [_jittered_tri2, None])
def test_delta_of_a_uniform_state_has_the_closed_form(gen):
    m = gen(7, 0.2, seed=5) if gen is not None else omesh.tri4(5)
    geo = oswe.geometry(m)
    brute = _brute_geometry(m)
    nc = m['cells'].shape[0]
    Wx, Wy = np.zeros(nc), np.zeros(nc)
    np.add.at(Wx, m['bcells'], -brute['bsx'])      # outward wall normals
    np.add.at(Wy, m['bcells'], -brute['bsy'])
    assert np.abs(Wx).max() > 0 and np.abs(Wy).max() > 0
    H, U, V, g = 1.3, 0.7, -0.4, 9.81
    s = oswe.SWE(m, geo, dt=1e-3, g=g)
    dh, duh, dvh = s.delta(torch.full((nc,), H, dtype=torch.float64),
                           torch.full((nc,), U, dtype=torch.float64),
                           torch.full((nc,), V, dtype=torch.float64))
    A = brute['area']
    np.testing.assert_allclose(dh.numpy(), -(U * Wx + V * Wy) / A, rtol=0, atol=1e-11)
    np.testing.assert_allclose(duh.numpy(), -(U * U / H * Wx + U * V / H * Wy) / A,
                               rtol=0, atol=1e-11)
    np.testing.assert_allclose(dvh.numpy(), -(U * V / H * Wx + V * V / H * Wy) / A,
                               rtol=0, atol=1e-11)


# ------------------------------------------------------------- D4 symmetry --
def _cell_keys(m, n):
    cent = m['points'][m['cells']].mean(axis=1)
    return np.rint(cent * 6 * n).astype(np.int64)   # tri4 centroids sit on d/6


def _symmetry_perm(m, n, kind):
    keys = _cell_keys(m, n)
    L = 6 * n
    X, Y = keys[:, 0], keys[:, 1]
    if kind == 'rot90':          # (x, y) -> (1 - y, x)
        TX, TY = L - Y, X
    else:                        # reflection (x, y) -> (1 - x, y)
        TX, TY = L - X, Y
    where = {(int(a), int(b)): c for c, (a, b) in enumerate(keys.tolist())}
    return np.array([where[(int(a), int(b))] for a, b in zip(TX, TY)])


def _transform_point(p, kind):
    x, y = p
    return (1 - y, x) if kind == 'rot90' else (1 - x, y)


def _transform_vec(u, v, kind):
    return (-v, u) if kind == 'rot90' else (-u, v)


def _equivariance_error(swe_cls, kind, n=8, steps=10):
    m = omesh.tri4(n)
    perm = _symmetry_perm(m, n, kind)
    x0 = (0.43, 0.61)
    a = swe_cls(m, oswe.geometry(m, centre=x0, width=200.0), dt=0.004, g=9.81)
    b = swe_cls(m, oswe.geometry(m, centre=_transform_point(x0, kind), width=200.0),
                dt=0.004, g=9.81)
    for _ in range(steps):
        a()
        b()
    tu, tv = _transform_vec(a.uh, a.vh, kind)
    err_h = float((b.h[perm] - a.h).abs().max())
    err_m = max(float((b.uh[perm] - tu).abs().max()),
                float((b.vh[perm] - tv).abs().max()))
    scale_m = float(torch.maximum(a.uh.abs(), a.vh.abs()).max())
    return err_h, err_m, scale_m


@pytest.mark.parametrize('kind', ['rot90', 'reflect_x'])
def test_swe_step_is_equivariant_under_the_mesh_symmetries(kind):
    err_h, err_m, scale_m = _equivariance_error(oswe.SWE, kind)
    assert scale_m > 1e-3                       # the flow is not trivial
    assert err_h <= 1e-13 and err_m <= 1e-13, (err_h, err_m)


class _SwappedPressure(oswe.SWE):
    """A planted mistake: the pressure term of delta_uh uses sy."""

    def delta(self, h, uh, vh):
        sx, sy = self.sx, self.sy
        dh, _, dvh = super().delta(h, uh, vh)
        self.sx, self.sy = sy, sx
        try:
            _, duh_bad, _ = super().delta(h, uh, vh)
        finally:
            self.sx, self.sy = sx, sy
        return dh, duh_bad, dvh


class _SwappedVelocity(oswe.SWE):
    """A planted mistake: u_f and v_f exchanged in the advection terms."""

    def delta(self, h, uh, vh):
        dh, _, _ = super().delta(h, uh, vh)
        _, duh_bad, dvh_bad = super().delta(h, vh, uh)
        return dh, dvh_bad, duh_bad


@pytest.mark.parametrize('bad', [_SwappedPressure, _SwappedVelocity])
def test_planted_transpositions_break_equivariance(bad):
    errs = [_equivariance_error(bad, k)[:2] for k in ('rot90', 'reflect_x')]
    assert max(max(e) for e in errs) > 1e-6, errs


# ------------------------------------------------------------- RK4 combine --
class _Linear(oswe.SWE):
    a = -3.0

    def delta(self, h, uh, vh):
        return self.a * h, self.a * uh, self.a * vh


class _ClassicalRK4(_Linear):
    def __call__(self):
        dt = self.dt
        d1 = self.delta(self.h, self.uh, self.vh)
        d2 = self.delta(*[u + 0.5 * dt * d for u, d in zip((self.h, self.uh, self.vh), d1)])
        d3 = self.delta(*[u + 0.5 * dt * d for u, d in zip((self.h, self.uh, self.vh), d2)])
        d4 = self.delta(*[u + dt * d for u, d in zip((self.h, self.uh, self.vh), d3)])
        for u, k in zip((self.h, self.uh, self.vh), range(3)):
            u[:] += dt / 6 * (d1[k] + 2 * d2[k] + 2 * d3[k] + d4[k])


def _amplification(cls):
    m = omesh.tri2(3)
    geo = oswe.geometry(m, centre=(0.5, 0.5), width=50.0)
    s = cls(m, geo, dt=0.05, g=9.81)
    s.uh[:] = torch.linspace(0.5, 1.5, s.nc, dtype=torch.float64)
    s.vh[:] = 0.25
    h0, uh0, vh0 = s.h.clone(), s.uh.clone(), s.vh.clone()
    s()
    return (s.h / h0, s.uh / uh0, s.vh / vh0), s.a * s.dt


def test_rk4_combine_has_the_reference_amplification_polynomial():
    (gh, gu, gv), z = _amplification(_Linear)
    want = 1 + 2 * z / 3 + z ** 2 / 3 + z ** 3 / 8 + z ** 4 / 24
    for g in (gh, gu, gv):
        np.testing.assert_allclose(g.numpy(), want, rtol=1e-14)
    (ch, _, _), _ = _amplification(_ClassicalRK4)
    classical = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    np.testing.assert_allclose(ch.numpy(), classical, rtol=1e-14)
    assert abs(classical - want) > 1e-3         # the check tells them apart
