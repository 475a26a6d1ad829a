"""Triangular mesh generators restated in numpy.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

tri4: csrc/triangular_mesh.cpp:10-142 -- 4 triangles per square around its
centre. Where the reference writes a shared grid point from several squares
(an OpenMP race in the reference), the value written last in sequential
square order is used; the reference built single-threaded (oracle/_ref) gives
exactly that.
tri2: 2 triangles per square (no reference counterpart; the configs'
"structured triangulation ... each cell split into 2 triangles").
"""
import numpy as np


def tri4(n: int):
    d = 1.0 / n
    W = n + 1
    nc, ne, npnt = 4 * n * n, 12 * n * n - 4 * n, W * W + n * n
    src = np.zeros(ne, np.int64)
    dst = np.zeros(ne, np.int64)
    cells = np.zeros((nc, 3), np.int64)
    points = np.zeros((npnt, 2), np.float64)
    bcells = np.zeros(4 * n, np.int64)
    bpoints = np.zeros((4 * n, 2), np.int64)
    for j in range(n):                      # triangular_mesh.cpp:31-129
        for i in range(n):
            nrect = i + j * n
            ncell = 4 * nrect
            p0 = j * W + i
            p1 = (j + 1) * W + i
            p2 = (j + 1) * W + i + 1
            p3 = j * W + i + 1
            p4 = W * W + n * j + i
            points[p0] = (i * d, j * d)
            points[p1] = (i * d, j * d + d)
            points[p2] = (i * d + d, j * d + d)
            points[p3] = (i * d + d, j * d)
            points[p4] = (i * d + d / 2, j * d + d / 2)
            cells[ncell] = (p0, p3, p4)
            cells[ncell + 1] = (p2, p4, p3)
            cells[ncell + 2] = (p2, p1, p4)
            cells[ncell + 3] = (p0, p4, p1)
            off = nrect * 8
            for k in range(4):
                a = ncell + k
                b = ncell if k == 3 else ncell + k + 1
                src[off + 2 * k], dst[off + 2 * k] = a, b
                src[off + 2 * k + 1], dst[off + 2 * k + 1] = b, a
            if j != 0:
                o = 8 * n * n + n * (j - 1) + i
                src[o], dst[o] = ncell, ncell - 4 * n + 2
            else:
                bcells[i] = ncell
                bpoints[i] = (p0, p3)
            if j != n - 1:
                o = 8 * n * n + n * (n - 1) + n * j + i
                src[o], dst[o] = ncell + 2, ncell + 4 * n
            else:
                bcells[n + i] = ncell + 2
                bpoints[n + i] = (p1, p2)
            if i != 0:
                o = 8 * n * n + 2 * n * (n - 1) + (n - 1) * j + i - 1
                src[o], dst[o] = ncell + 3, ncell - 3
            else:
                bcells[2 * n + j] = ncell + 3
                bpoints[2 * n + j] = (p0, p1)
            if i != n - 1:
                o = 8 * n * n + 3 * n * (n - 1) + (n - 1) * j + i
                src[o], dst[o] = ncell + 1, ncell + 7
            else:
                bcells[3 * n + j] = ncell + 1
                bpoints[3 * n + j] = (p2, p3)
    return dict(src=src, dst=dst, cells=cells, bcells=bcells,
                bpoints=bpoints, points=points)


def tri2(n: int):
    """2 triangles per square, diagonal from (i, j) to (i+1, j+1); the same six
    arrays as tri4 (schema of tutorial/create_triangular_mesh.py:29-35)."""
    W = n + 1
    I, J = np.meshgrid(np.arange(W), np.arange(W))
    points = np.stack([I.ravel() / n, J.ravel() / n], axis=1).astype(np.float64)
    cells, src, dst, bcells, bpoints = [], [], [], [], []
    for j in range(n):
        for i in range(n):
            p00, p10 = j * W + i, j * W + i + 1
            p01, p11 = (j + 1) * W + i, (j + 1) * W + i + 1
            cells.append((p00, p10, p11))
            cells.append((p00, p11, p01))
    cells = np.array(cells, np.int64)

    def c(i, j, k):
        return 2 * (j * n + i) + k
    for j in range(n):
        for i in range(n):
            src += [c(i, j, 0), c(i, j, 1)]
            dst += [c(i, j, 1), c(i, j, 0)]
    for j in range(n):
        for i in range(n - 1):
            src += [c(i, j, 0), c(i + 1, j, 1)]
            dst += [c(i + 1, j, 1), c(i, j, 0)]
    for j in range(n - 1):
        for i in range(n):
            src += [c(i, j, 1), c(i, j + 1, 0)]
            dst += [c(i, j + 1, 0), c(i, j, 1)]
    for i in range(n):
        bcells.append(c(i, 0, 0))
        bpoints.append((i, i + 1))
    for i in range(n):
        bcells.append(c(i, n - 1, 1))
        bpoints.append(((n) * W + i, n * W + i + 1))
    for j in range(n):
        bcells.append(c(0, j, 1))
        bpoints.append((j * W, (j + 1) * W))
    for j in range(n):
        bcells.append(c(n - 1, j, 0))
        bpoints.append((j * W + n, (j + 1) * W + n))
    return dict(src=np.array(src, np.int64), dst=np.array(dst, np.int64),
                cells=cells, bcells=np.array(bcells, np.int64),
                bpoints=np.array(bpoints, np.int64), points=points)


def dual_edges_from_cells(cells: np.ndarray):
    """Brute force: the set of ordered cell pairs sharing two vertices."""
    faces = {}
    for t, tri in enumerate(cells.tolist()):
        for a, b in ((tri[0], tri[1]), (tri[1], tri[2]), (tri[2], tri[0])):
            faces.setdefault((min(a, b), max(a, b)), []).append(t)
    pairs, boundary = set(), set()
    for f, ts in faces.items():
        if len(ts) == 2:
            pairs.add((ts[0], ts[1]))
            pairs.add((ts[1], ts[0]))
        else:
            boundary.add((ts[0], f))
    return pairs, boundary
