"""Smoothed-aggregation algebraic multigrid (Vanek, Mandel and Brezina,
Computing 56, 1996), restated in plain Python loops and scipy.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

No reference counterpart: the reference's preconditioner slot
(easier/numeric/preconditioner/__init__.py:4) is empty and its solvers take
any callable M (numeric/solver/cg.py:86). This restates the hierarchy that
the multilevel preconditioner builds for an operator on any mesh:

* strength: j is a strong neighbour of i (j != i) when a_ij != 0 and
  a_ij^2 >= theta^2 |a_ii a_jj|;
* aggregates, vertices visited in index order: pass 1 -- a vertex with strong
  neighbours, none of them aggregated, and itself free, forms an aggregate
  with all its strong neighbours; pass 2 -- a vertex still free joins the
  pass-1 aggregate of its strongest aggregated strong neighbour (the first
  of equals); pass 3 -- a vertex still free with strong neighbours forms an
  aggregate with those still free. Vertices without strong neighbours stay
  out (-1);
* P = (I - omega D^-1 A) P_tent, P_tent[i, agg(i)] = 1, omega = 4 / (3 rho),
  rho = max_i (|a_ii| + sum_j |a_ij|) / |a_ii| (Gershgorin bound of D^-1 A);
* Ac = P^T A P; theta halves on every coarser level; coarsening stops at
  coarse_rows rows (dense inverse) or when a level keeps over 90 % of its
  rows;
* V(1,1) cycle with damped Jacobi of weight omega 2 / rho on each level.

Pinned in tests/test_amg_host.py: the aggregates by their defining
properties (pass-1 aggregates are a root with exactly its strong
neighbours, every aggregate is connected in the strength graph, every
vertex with a strong neighbour is aggregated once), P and Ac by scipy's
sparse products.
"""
import types

import numpy as np
import scipy.sparse as sp


def strength(A, theta):
    """Boolean CSR of the strong couplings of A (scipy CSR, diagonal stored)."""
    A = sp.csr_matrix(A)
    d = A.diagonal()
    coo = A.tocoo()
    keep = (coo.row != coo.col) & (coo.data != 0) & \
        (coo.data ** 2 >= theta ** 2 * np.abs(d[coo.row] * d[coo.col]))
    S = sp.csr_matrix((np.ones(keep.sum()), (coo.row[keep], coo.col[keep])), shape=A.shape)
    S.sort_indices()
    return S


def aggregate(A, theta):
    """Aggregate index per vertex (-1: not aggregated) and the pass (1, 2 or
    3) that placed it (0: none)."""
    A = sp.csr_matrix(A)
    S = strength(A, theta)
    d = A.diagonal()
    n = A.shape[0]
    agg = [-1] * n
    how = [0] * n
    nbrs = [S.indices[S.indptr[i]:S.indptr[i + 1]].tolist() for i in range(n)]
    na = 0
    for i in range(n):
        if not nbrs[i] or agg[i] >= 0 or any(agg[j] >= 0 for j in nbrs[i]):
            continue
        for j in [i] + nbrs[i]:
            agg[j] = na
            how[j] = 1
        na += 1
    Ad = A.todok() if n <= 4000 else None
    joined = {}
    for i in range(n):
        if agg[i] >= 0 or not nbrs[i]:
            continue
        best, pick = -1.0, -1
        for j in nbrs[i]:
            if agg[j] < 0:
                continue
            a = Ad[i, j] if Ad is not None else A[i, j]
            s = abs(a) / np.sqrt(abs(d[i] * d[j]))
            if s > best:
                best, pick = s, agg[j]
        if pick >= 0:
            joined[i] = pick
    for i, a in joined.items():
        agg[i] = a
        how[i] = 2
    for i in range(n):
        if agg[i] >= 0 or not nbrs[i]:
            continue
        agg[i] = na
        how[i] = 3
        for j in nbrs[i]:
            if agg[j] < 0:
                agg[j] = na
                how[j] = 3
        na += 1
    return np.array(agg, np.int64), np.array(how, np.int64)


def level(A, theta):
    """One level: dict(agg, how, rho, P, Ac) (P, Ac scipy CSR)."""
    A = sp.csr_matrix(A)
    n = A.shape[0]
    d = A.diagonal()
    rho = float(np.max(np.abs(A).sum(axis=1).A1 / np.abs(d)))
    agg, how = aggregate(A, theta)
    nc = int(agg.max()) + 1 if np.any(agg >= 0) else 0
    ok = agg >= 0
    T = sp.csr_matrix((np.ones(ok.sum()), (np.nonzero(ok)[0], agg[ok])), shape=(n, nc))
    P = ((sp.eye(n) - (4.0 / (3.0 * rho)) * sp.diags(1.0 / d) @ A) @ T).tocsr()
    Ac = (P.T @ A @ P).tocsr()
    return dict(agg=agg, how=how, rho=rho, P=P, Ac=Ac, nc=nc)


def hierarchy(A, theta=0.08, coarse_rows=1024, max_levels=25):
    """Levels (namespaces with A, d, P, rho, Ainv) down to the dense one."""
    A = sp.csr_matrix(A)
    levels = [types.SimpleNamespace(A=A, d=A.diagonal(), P=None, rho=None, Ainv=None)]
    while levels[-1].A.shape[0] > coarse_rows and len(levels) < max_levels:
        L = levels[-1]
        lv = level(L.A, theta * 0.5 ** (len(levels) - 1))
        if lv['nc'] == 0 or lv['nc'] > 0.9 * L.A.shape[0]:
            break
        L.P, L.rho = lv['P'], lv['rho']
        levels.append(types.SimpleNamespace(A=lv['Ac'], d=lv['Ac'].diagonal(), P=None,
                                            rho=None, Ainv=None))
    last = levels[-1]
    last.Ainv = np.linalg.inv(last.A.toarray())
    return levels


def vcycle(levels, r, omega=0.8, l=0):
    """z ~ A^{-1} r: z = w D^-1 r; t = r - A z; z += P V(P^T t);
    z += w D^-1 (r - A z), w = omega 2 / rho on level l."""
    L = levels[l]
    if L.Ainv is not None:
        return L.Ainv @ r
    w = omega * 2.0 / L.rho
    z = w * r / L.d
    t = r - L.A @ z
    zc = vcycle(levels, L.P.T @ t, omega, l + 1)
    z = z + L.P @ zc
    return z + w * (r - L.A @ z) / L.d
