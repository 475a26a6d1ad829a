"""Poisson problems of the benchmark configurations: P1 finite elements on
the 2-triangle mesh, zero Dirichlet boundary, b = M f, optional
variable coefficient kappa = exp(Gaussian random field).

PoissonP1 is the program the reference's Poisson tutorial builds
(tutorial/poisson/poisson_main.py:16-53: x, b, Ac/Av, Af/Ae and
A = Linsys(Av, Ae, Selector(col), Reducer(row))) with the P1 operator in place
of the tutorial's finite-volume one.
"""
import ctypes
import math
import os

import numpy as np
import torch

import easier_amd as esr
from .. import _native as nat
from ..numeric import Linsys
from ..numeric.preconditioner import Jacobi
from . import mesh as wmesh


def p1_system(mesh, kappa_cell=None, f=None):
    """Assemble the P1 system on host; returns numpy arrays row, col (int64,
    sorted by row then col), Ae, Av, b (float64) and the boundary mask."""
    pts = np.ascontiguousarray(mesh['points'], dtype=np.float64)
    cells = np.ascontiguousarray(mesh['cells'], dtype=np.int64)
    npnt, nc = pts.shape[0], cells.shape[0]
    nnz = ctypes.c_int64()
    nat.call('esr_p1_nnz', npnt, nc, cells.ctypes.data, ctypes.byref(nnz))
    m = nnz.value
    row = np.empty(m, np.int64)
    col = np.empty(m, np.int64)
    Ae = np.empty(m, np.float64)
    Av = np.empty(npnt, np.float64)
    b = np.empty(npnt, np.float64)
    bnd = np.empty(npnt, np.uint8)
    kap = None if kappa_cell is None else np.ascontiguousarray(kappa_cell, np.float64)
    ff = None if f is None else np.ascontiguousarray(f, np.float64)
    nat.call('esr_p1_assemble', npnt, nc, cells.ctypes.data, pts.ctypes.data,
             None if kap is None else kap.ctypes.data,
             None if ff is None else ff.ctypes.data, m, row.ctypes.data,
             col.ctypes.data, Ae.ctypes.data, Av.ctypes.data, b.ctypes.data,
             bnd.ctypes.data)
    return dict(row=row, col=col, Ae=Ae, Av=Av, b=b, boundary=bnd.astype(bool))


def load_vectors(mesh, fs, boundary):
    """b = M f for each row f of fs (k x vertices): the consistent P1 mass
    matrix (|T|/12 (1 + [i == j]) per triangle), zero on boundary vertices --
    the load vector of p1_system, contributions added in the same triangle
    order (so the same bits) -- for many right-hand sides of one mesh without
    re-assembling the operator."""
    pts = np.asarray(mesh['points'], np.float64)
    c = np.asarray(mesh['cells'], np.int64)
    xa, ya = pts[c[:, 0], 0], pts[c[:, 0], 1]
    xb, yb = pts[c[:, 1], 0], pts[c[:, 1], 1]
    xc, yc = pts[c[:, 2], 0], pts[c[:, 2], 1]
    cross = (xb - xa) * (yc - ya) - (xc - xa) * (yb - ya)
    m = (0.5 * np.abs(cross)) / 12.0
    bnd = np.asarray(boundary, bool)
    out = np.empty((len(fs), pts.shape[0]))
    flat = c.ravel()

    def one(k):
        f = np.asarray(fs[k], np.float64)
        fv = f[c]                                   # (triangles, 3)
        contrib = np.empty_like(fv)
        for i in range(3):
            contrib[:, i] = m * (2.0 * fv[:, i] + fv[:, (i + 1) % 3] + fv[:, (i + 2) % 3])
        out[k] = np.bincount(flat, weights=contrib.ravel(), minlength=pts.shape[0])
        out[k, bnd] = 0.0

    if len(fs) > 1:
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(min(len(fs), os.cpu_count() or 1, 16)) as ex:
            list(ex.map(one, range(len(fs))))
    elif fs:
        one(0)
    return out


def gaussian_random_field(n, seed, ell=0.05, variance=1.0):
    """A stationary Gaussian field with covariance
    variance * exp(-|x - y|^2 / (2 ell^2)) on the (n+1) x (n+1) vertex grid of
    the unit square (spacing 1/n), synthesised exactly by circulant embedding:
    z = ifft2(sqrt(lambda) * fft2(w)), lambda = fft2 of the covariance on a
    periodic grid large enough that the kernel has decayed (>= 6 ell of margin),
    w white noise from numpy's PCG64 seeded with `seed`."""
    from scipy import fft as sfft
    h = 1.0 / n
    need = (n + 1) + int(math.ceil(6 * ell / h))
    N = sfft.next_fast_len(need)
    k = np.arange(N)
    d = np.minimum(k, N - k) * h
    c1 = np.exp(-d ** 2 / (2 * ell ** 2))
    lam = np.real(sfft.fft(c1))
    lam2 = variance * np.outer(lam, lam)          # separable kernel
    lam2 = np.maximum(lam2, 0.0)
    rng = np.random.Generator(np.random.PCG64(seed))
    w = rng.standard_normal((N, N))
    z = np.real(sfft.ifft2(np.sqrt(lam2) * sfft.fft2(w)))
    return z[:n + 1, :n + 1].copy()


def kappa_cells(mesh, log_kappa_vertices):
    """kappa_T = exp(mean of log kappa over the triangle's vertices)."""
    g = np.asarray(log_kappa_vertices, dtype=np.float64).ravel()
    c = mesh['cells']
    return np.exp((g[c[:, 0]] + g[c[:, 1]] + g[c[:, 2]]) / 3.0)


def rhs_uniform(npnt, seed):
    """f ~ U[-1, 1] per vertex from torch.Generator('cpu').manual_seed(seed)."""
    gen = torch.Generator().manual_seed(seed)
    return (torch.rand(npnt, generator=gen, dtype=torch.float64) * 2 - 1).numpy()


class PoissonP1(esr.Module):
    """x, b, Av, Ae, A = Linsys(Av, Ae, Selector(col), Reducer(row)) and the
    Jacobi preconditioner M(r) = r / Av."""

    def __init__(self, system, dtype=torch.float64, x=None):
        super().__init__()
        n = system['Av'].shape[0]
        self.reducer = esr.Reducer(torch.from_numpy(system['row']), n)
        self.selector = esr.Selector(torch.from_numpy(system['col']))
        self.x = esr.Tensor(esr.zeros((n,), dtype=dtype), mode='partition') \
            if x is None else x
        self.b = esr.Tensor(torch.from_numpy(system['b']).to(dtype), mode='partition')
        self.Av = esr.Tensor(torch.from_numpy(system['Av']).to(dtype), mode='partition')
        self.Ae = esr.Tensor(torch.from_numpy(system['Ae']).to(dtype), mode='partition')
        self.A = Linsys(self.Av, self.Ae, self.selector, self.reducer)
        self.M = Jacobi(self.Av)


def config1_system(n=100, seed=0):
    """P1 Poisson, kappa = 1, on the n x n 2-triangle mesh; f ~ U[-1,1]."""
    mesh = wmesh.tri2(n)
    f = rhs_uniform(mesh['points'].shape[0], seed)
    return mesh, p1_system(mesh, None, f)


def varcoef_system(n, seed, jitter=0.0, ell=0.05):
    """Variable-coefficient P1 Poisson: log kappa a seeded Gaussian random
    field (ell = 0.05, variance 1), f ~ U[-1, 1] (seed + 1)."""
    mesh = wmesh.tri2(n, jitter=jitter, seed=seed)
    logk = gaussian_random_field(n, seed, ell=ell)
    kap = kappa_cells(mesh, logk)
    f = rhs_uniform(mesh['points'].shape[0], seed + 1)
    return mesh, p1_system(mesh, kap, f), kap
