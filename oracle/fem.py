"""P1 finite elements restated in numpy.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The reference has no P1 assembler (its Poisson tutorial is a cell-centred
finite-volume scheme, tutorial/poisson/assemble_poisson.py:160-230). This
restatement is pinned by closed forms in tests/test_oracle_fem.py: on the
uniform 2-triangle mesh with kappa = 1 the operator is the 5-point stencil
(diagonal-edge weights cot 90 deg = 0), the mass load of f = 1 is h^2 per
interior vertex, rows of kappa-constant operators sum to zero away from the
boundary, and the matrix is symmetric.
"""
import numpy as np


def assemble(mesh, kappa_cell=None, f=None):
    pts = mesh['points']
    cells = mesh['cells']
    npnt, nc = pts.shape[0], cells.shape[0]
    kap = np.ones(nc) if kappa_cell is None else np.asarray(kappa_cell, np.float64)
    a, b_, c = cells[:, 0], cells[:, 1], cells[:, 2]
    xa, ya = pts[a, 0], pts[a, 1]
    xb, yb = pts[b_, 0], pts[b_, 1]
    xc, yc = pts[c, 0], pts[c, 1]
    cross = (xb - xa) * (yc - ya) - (xc - xa) * (yb - ya)
    area = 0.5 * np.abs(cross)
    s = kap / (4.0 * area)
    ex = np.stack([xc - xb, xa - xc, xb - xa], axis=1)
    ey = np.stack([yc - yb, ya - yc, yb - ya], axis=1)
    # unique undirected edges and their multiplicity
    e = np.concatenate([cells[:, [0, 1]], cells[:, [1, 2]], cells[:, [2, 0]]])
    e = np.sort(e, axis=1)
    uniq, counts = np.unique(e, axis=0, return_counts=True)
    boundary = np.zeros(npnt, bool)
    boundary[uniq[counts == 1].ravel()] = True
    row = np.concatenate([uniq[:, 0], uniq[:, 1]])
    col = np.concatenate([uniq[:, 1], uniq[:, 0]])
    order = np.lexsort((col, row))
    row, col = row[order], col[order]
    key = {(int(r), int(cc)): k for k, (r, cc) in enumerate(zip(row, col))}
    Ae = np.zeros(row.shape[0])
    Av = np.zeros(npnt)
    bvec = np.zeros(npnt)
    v = cells
    for t in range(nc):                       # triangle order, as assembled
        for i in range(3):
            Av[v[t, i]] += s[t] * (ex[t, i] * ex[t, i] + ey[t, i] * ey[t, i])
            for j in range(3):
                if i != j:
                    Ae[key[(int(v[t, i]), int(v[t, j]))]] += \
                        s[t] * (ex[t, i] * ex[t, j] + ey[t, i] * ey[t, j])
        if f is not None:
            m = area[t] / 12.0
            for i in range(3):
                bvec[v[t, i]] += m * (2.0 * f[v[t, i]] + f[v[t, (i + 1) % 3]]
                                      + f[v[t, (i + 2) % 3]])
    Av[boundary] = 1.0
    bvec[boundary] = 0.0
    Ae[boundary[row] | boundary[col]] = 0.0
    return dict(row=row, col=col, Ae=Ae, Av=Av, b=bvec, boundary=boundary)
