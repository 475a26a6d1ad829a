"""Structured triangular meshes from the native generator
(csrc/triangular_mesh.cpp:10-142 for kind 4; 2 triangles per square for
kind 2). Arrays follow tutorial/create_triangular_mesh.py:29-35:
src, dst (ne,), cells (nc, 3), bcells (nbc,), bpoints (nbc, 2) int64 and
points (np, 2) float64."""
import ctypes

import numpy as np

from .. import _native as nat


def _sizes(kind, n):
    sz = [ctypes.c_int64() for _ in range(4)]
    nat.call('esr_mesh_sizes', kind, n, *[ctypes.byref(s) for s in sz])
    return [s.value for s in sz]


def _alloc(kind, n):
    npnt, nc, ne, nbc = _sizes(kind, n)
    return dict(src=np.empty(ne, np.int64), dst=np.empty(ne, np.int64),
                cells=np.empty((nc, 3), np.int64),
                bcells=np.empty(nbc, np.int64),
                bpoints=np.empty((nbc, 2), np.int64),
                points=np.empty((npnt, 2), np.float64))


_ORDER = ('src', 'dst', 'cells', 'bcells', 'bpoints', 'points')


def tri4(n: int):
    """4 triangles per square around its centre (the reference's mesh)."""
    m = _alloc(4, n)
    nat.call('esr_mesh_tri4', n, *[m[k].ctypes.data for k in _ORDER])
    return m


def tri2(n: int, jitter: float = 0.0, seed: int = 0):
    """2 triangles per square; interior points optionally jittered by
    uniform(-jitter, jitter) * (1/n) in x and y."""
    m = _alloc(2, n)
    nat.call('esr_mesh_tri2', n, float(jitter), int(seed),
             *[m[k].ctypes.data for k in _ORDER])
    return m


def save_h5(mesh, path: str) -> None:
    """Write a mesh in the schema of tutorial/create_triangular_mesh.py:29-35
    (datasets src, dst, cells, bcells, points, bpoints) with the package's
    HDF5 writer; h5py and the reference's loaders read the file."""
    from ..core import h5
    h5.write(path, {k: np.asarray(mesh[k]) for k in ('src', 'dst', 'cells', 'bcells',
                                                      'points', 'bpoints')})


def load_h5(path: str):
    """A mesh file of that schema (as create_triangular_mesh.py writes it with
    h5py) as the dict tri4/tri2 return: int64 indices, float64 points."""
    from ..core import h5
    out = {}
    with h5.H5Reader(path) as r:
        for k in _ORDER:
            a = r.read(k)
            out[k] = np.ascontiguousarray(a, np.float64 if k == 'points' else np.int64)
    return out
