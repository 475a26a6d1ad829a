"""Geometric multigrid on the 2-triangle mesh, restated in numpy/scipy.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

No reference counterpart: the reference's preconditioner slot
(easier/numeric/preconditioner/__init__.py:4) is empty and its solvers take
any callable M (numeric/solver/cg.py:86). This restates the hierarchy the
multigrid preconditioner builds on the GPU:

* grids: the (n+1) x (n+1) vertex grid of the 2-triangle mesh (vertex
  (I, J) at lexicographic position J (n+1) + I, each square split along its
  SW-NE diagonal, as workloads/mesh.tri2 builds it); level l+1 has n/2 cells
  per side and is kept in lexicographic order;
* prolongation P: P1 interpolation between nested meshes -- a fine vertex at
  an even-even position takes its coarse parent's value, one on a horizontal,
  vertical or SW-NE-diagonal coarse edge the mean of the edge's two ends.
  Columns of coarse boundary vertices are dropped (zero Dirichlet values);
* coarse operator: the Galerkin product P^T A P, with identity rows for the
  coarse boundary vertices; its pattern is the 7-point stencil of the
  coarse 2-triangle mesh;
* V(1,1) cycle with damped-Jacobi smoothing and an exact (dense) solve on
  the coarsest grid.

Pinned in tests/test_oracle_multigrid.py: P against barycentric evaluation
of coarse P1 functions at fine vertices (fixes the diagonal direction); the
Galerkin operator of a constant coefficient against the coarse-mesh P1
assembly (nested P1 spaces: P^T A_h P = A_2h exactly); the restriction
against scipy's P^T product; the pattern against the P1 assembly's pattern;
the V-cycle's symmetry and its contraction on the model problem.
"""
import numpy as np
import scipy.sparse as sp

# coarse-stencil slots in ascending column order of a lexicographic row:
# SW, S, W, E, N, NE (offsets (dx, dy))
STENCIL = ((-1, -1), (0, -1), (-1, 0), (1, 0), (0, 1), (1, 1))
# children of a coarse vertex (offsets on the fine grid) with P weights, in
# the order the restriction folds them
CHILDREN = ((0, 0, 1.0), (-1, 0, 0.5), (1, 0, 0.5), (0, -1, 0.5), (0, 1, 0.5),
            (-1, -1, 0.5), (1, 1, 0.5))


def grid_positions(points, n):
    """Lexicographic grid position J (n+1) + I of each vertex, from its
    coordinates (I = rint(x n), J = rint(y n); jitter below half a cell keeps
    the rounding exact). Raises unless it is a bijection."""
    pts = np.asarray(points, np.float64)
    I = np.rint(pts[:, 0] * n).astype(np.int64)
    J = np.rint(pts[:, 1] * n).astype(np.int64)
    W = n + 1
    if pts.shape[0] != W * W or I.min() < 0 or J.min() < 0 or I.max() > n or J.max() > n:
        raise ValueError('points are not the vertices of an n x n grid')
    pos = J * W + I
    if np.unique(pos).shape[0] != W * W:
        raise ValueError('two vertices map to one grid position')
    return pos


def interior(n):
    W = n + 1
    I, J = np.meshgrid(np.arange(W), np.arange(W))
    return ((I > 0) & (I < n) & (J > 0) & (J < n)).reshape(-1)


def parents(I, J):
    """Coarse parents (a, b) and weights of fine vertex (I, J)."""
    if I % 2 == 0 and J % 2 == 0:
        return [(I // 2, J // 2, 1.0)]
    if J % 2 == 0:
        return [((I - 1) // 2, J // 2, 0.5), ((I + 1) // 2, J // 2, 0.5)]
    if I % 2 == 0:
        return [(I // 2, (J - 1) // 2, 0.5), (I // 2, (J + 1) // 2, 0.5)]
    return [((I - 1) // 2, (J - 1) // 2, 0.5), ((I + 1) // 2, (J + 1) // 2, 0.5)]


def prolongation(nf, fpos=None, keep_boundary=False):
    """P (fine vertices x coarse vertices) as scipy CSR; fine rows in the
    order fpos gives (fine index -> lexicographic position; None = lex)."""
    if nf % 2:
        raise ValueError('odd grids are not coarsened')
    nc = nf // 2
    Wf, Wc = nf + 1, nc + 1
    nvf = Wf * Wf
    pos = np.arange(nvf) if fpos is None else np.asarray(fpos)
    rows, cols, vals = [], [], []
    for i in range(nvf):
        I, J = int(pos[i]) % Wf, int(pos[i]) // Wf
        for a, b, w in parents(I, J):
            if keep_boundary or (0 < a < nc and 0 < b < nc):
                rows.append(i)
                cols.append(b * Wc + a)
                vals.append(w)
    return sp.csr_matrix((vals, (rows, cols)), shape=(nvf, Wc * Wc))


def galerkin(A, P, nc):
    """P^T A P with identity rows/columns for the coarse boundary vertices."""
    Ac = (P.T @ A @ P).tocsr()
    bnd = ~interior(nc)
    Ac = Ac.tolil()
    for c in np.nonzero(bnd)[0]:
        Ac[c, c] = 1.0
    return Ac.tocsr()


def pattern(nc):
    """rowptr, col of the coarse 2-triangle mesh's vertex graph in
    lexicographic order (neighbours ascending: SW, S, W, E, N, NE)."""
    W = nc + 1
    rowptr, col = [0], []
    for y in range(W):
        for x in range(W):
            for dx, dy in STENCIL:
                if 0 <= x + dx < W and 0 <= y + dy < W:
                    col.append((y + dy) * W + x + dx)
            rowptr.append(len(col))
    return np.asarray(rowptr, np.int64), np.asarray(col, np.int64)


def linsys_csr(Ac, nc):
    """(rowptr, col, Ae, Av) of a coarse operator on pattern(nc): the
    off-diagonal values in the pattern's slots, the diagonal apart."""
    rowptr, col = pattern(nc)
    row = np.repeat(np.arange(rowptr.shape[0] - 1), np.diff(rowptr))
    Ae = np.asarray(Ac[row, col]).reshape(-1)
    Av = Ac.diagonal().copy()
    return rowptr, col, Ae, Av


def restrict(t, nf, fidx=None):
    """r_c = P^T t in the kernel's fold order: for an interior coarse vertex
    t[child 0] + 0.5 * (((((t_W + t_E) + t_S) + t_N) + t_SW) + t_NE); zero on
    the coarse boundary. fidx: lexicographic position -> fine index."""
    nc = nf // 2
    Wf, Wc = nf + 1, nc + 1
    tl = np.asarray(t)[np.arange(Wf * Wf) if fidx is None else np.asarray(fidx)]
    T = tl.reshape(Wf, Wf)                       # [J, I]
    out = np.zeros((Wc, Wc), dtype=tl.dtype)
    a = np.arange(1, nc)
    I0 = 2 * a
    g = lambda dx, dy: T[np.ix_(I0 + dy, I0 + dx)]
    acc = g(-1, 0)
    acc = acc + g(1, 0)
    acc = acc + g(0, -1)
    acc = acc + g(0, 1)
    acc = acc + g(-1, -1)
    acc = acc + g(1, 1)
    out[1:nc, 1:nc] = g(0, 0) + 0.5 * acc
    return out.reshape(-1)


def prolong_values(zc, nf, fpos=None):
    """(P zc) per fine vertex in the kernel's order: the parent's value, or
    0.5 * (first + second) for two parents; coarse boundary values read as 0."""
    nc = nf // 2
    Wf, Wc = nf + 1, nc + 1
    zcm = np.asarray(zc).reshape(Wc, Wc).copy()
    zcm[0, :] = 0
    zcm[-1, :] = 0
    zcm[:, 0] = 0
    zcm[:, -1] = 0
    pos = np.arange(Wf * Wf) if fpos is None else np.asarray(fpos)
    I, J = pos % Wf, pos // Wf
    v = np.empty(pos.shape[0], dtype=zcm.dtype)
    ee = (I % 2 == 0) & (J % 2 == 0)
    oe = (I % 2 == 1) & (J % 2 == 0)
    eo = (I % 2 == 0) & (J % 2 == 1)
    oo = (I % 2 == 1) & (J % 2 == 1)
    v[ee] = zcm[J[ee] // 2, I[ee] // 2]
    v[oe] = 0.5 * (zcm[J[oe] // 2, (I[oe] - 1) // 2] + zcm[J[oe] // 2, (I[oe] + 1) // 2])
    v[eo] = 0.5 * (zcm[(J[eo] - 1) // 2, I[eo] // 2] + zcm[(J[eo] + 1) // 2, I[eo] // 2])
    v[oo] = 0.5 * (zcm[(J[oo] - 1) // 2, (I[oo] - 1) // 2] +
                   zcm[(J[oo] + 1) // 2, (I[oo] + 1) // 2])
    return v


class Level:
    """One grid of the hierarchy: scipy operator A, grid size n, and for the
    finest level the fine index <-> lexicographic position maps."""

    def __init__(self, A, n, fpos=None):
        self.A, self.n = A.tocsr(), n
        self.d = self.A.diagonal().copy()
        self.fpos = None if fpos is None else np.asarray(fpos)
        self.fidx = None
        if fpos is not None:
            self.fidx = np.empty_like(self.fpos)
            self.fidx[self.fpos] = np.arange(self.fpos.shape[0])
        self.P = None
        self.Ainv = None


def hierarchy(A, n, fpos=None, coarsest=16):
    """Levels from the fine operator A (vertex order fpos) down to the first
    grid with n <= coarsest or n odd (solved densely)."""
    levels = [Level(A, n, fpos)]
    while levels[-1].n % 2 == 0 and levels[-1].n > coarsest:
        L = levels[-1]
        L.P = prolongation(L.n, L.fpos)
        Ac = galerkin(L.A, L.P, L.n // 2)
        levels.append(Level(Ac, L.n // 2))
    levels[-1].Ainv = np.linalg.inv(levels[-1].A.toarray())
    return levels


def vcycle(levels, r, omega=0.8, l=0, omega_pre=None):
    """z ~ A^{-1} r by one V(1,1) cycle: z = w' D^-1 r; t = r - A z;
    z += P V(P^T t); z += w D^-1 (r - A z), w' = omega_pre (default w) on
    every level."""
    L = levels[l]
    if L.Ainv is not None:
        return L.Ainv @ r
    wp = omega if omega_pre is None else omega_pre
    z = wp * r / L.d
    t = r - L.A @ z
    zc = vcycle(levels, L.P.T @ t, omega, l + 1, omega_pre)
    z = z + L.P @ zc
    return z + omega * (r - L.A @ z) / L.d
