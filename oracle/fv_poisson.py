# Portions of this file reproduce EASIER (jack-pan-ai/EASIER),
# Copyright (c) Microsoft Corporation, MIT License; the licence text is in
# easier_amd/THIRD_PARTY_NOTICES.txt.
"""Cell-centred finite-volume Poisson assembly restated with torch CPU
operators.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

assemble(): tutorial/poisson/assemble_poisson.py:160-230
(PoissonInitializer.forward) on an in-memory mesh; the index arrays the
tutorial's PoissonMeshComponentsCollector (:62-77) writes to the HDF5 file
are formed directly (cells[src][:, i], cells[dst][:, i], cells[:, i],
bpoints[:, i]). Returns Af (per directed edge), Ac, b, rho (per cell) and
the cell centroids, which tutorial/poisson/poisson_main.py:16-53 feeds to
Linsys(Ac, Af, Selector(dst), Reducer(src)).

Pinned by tests/test_oracle_fv_poisson.py: closed forms on the uniform
4-triangle mesh (Af = 3/2 across a half-diagonal, 3 across a square side,
Ac = -6 inside, -9 at a wall), a brute-force construction from each cell
pair's shared vertices on a jittered mesh, symmetry of the operator, and
its annihilation of linear functions on interior cells of the uniform mesh.
"""
import torch

from .ops import reducer, selector
from .swe import get_face_norm


def assemble(mesh, centre=(0.5, 0.5)):
    src = torch.as_tensor(mesh['src'])
    dst = torch.as_tensor(mesh['dst'])
    cells = torch.as_tensor(mesh['cells'])
    points = torch.as_tensor(mesh['points'], dtype=torch.float64)
    bcells = torch.as_tensor(mesh['bcells'])
    bpoints = torch.as_tensor(mesh['bpoints'])
    nc = cells.shape[0]
    src_p = selector(cells, src)
    dst_p = selector(cells, dst)
    sp = [selector(points, src_p[:, i]) for i in range(3)]
    dp = [selector(points, dst_p[:, i]) for i in range(3)]
    src_cent = (sp[0] + sp[1] + sp[2]) / 3.
    dst_cent = (dp[0] + dp[1] + dp[2]) / 3.
    dist = dst_cent - src_cent
    n01 = get_face_norm(sp[2], sp[0], sp[1])
    n12 = get_face_norm(sp[0], sp[1], sp[2])
    n20 = get_face_norm(sp[1], sp[2], sp[0])
    n01_ = get_face_norm(dst_cent, sp[0], sp[1])
    n12_ = get_face_norm(dst_cent, sp[1], sp[2])
    n20_ = get_face_norm(dst_cent, sp[2], sp[0])
    cond = (n01[0] * n01_[0] + n01[1] * n01_[1]) < 0
    nx = torch.where(cond, n01[0], 0.)
    ny = torch.where(cond, n01[1], 0.)
    cond = (n12[0] * n12_[0] + n12[1] * n12_[1]) < 0
    nx = torch.where(cond, n12[0], nx)
    ny = torch.where(cond, n12[1], ny)
    cond = (n20[0] * n20_[0] + n20[1] * n20_[1]) < 0
    nx = torch.where(cond, n20[0], nx)
    ny = torch.where(cond, n20[1], ny)
    dist = dist / (dist ** 2).sum(dim=1, keepdim=True)
    Af = dist[:, 0] * nx + dist[:, 1] * ny
    Ac = -reducer(Af, src, nc)
    p0 = selector(points, cells[:, 0])
    p1 = selector(points, cells[:, 1])
    p2 = selector(points, cells[:, 2])
    x0, y0 = p0[:, 0], p0[:, 1]
    x1, y1 = p1[:, 0], p1[:, 1]
    x2, y2 = p2[:, 0], p2[:, 1]
    area = 0.5 * torch.abs(x0 * (y1 - y2) + x1 * (y2 - y0) + x2 * (y0 - y1))
    centroid = (p0 + p1 + p2) / 3.
    c = torch.tensor([[float(centre[0]), float(centre[1])]], dtype=torch.float64)
    rho = torch.exp(-0.5 * 400 * ((centroid - c) ** 2).sum(1))
    b = rho * area
    b_p0 = selector(points, bpoints[:, 0])
    b_p1 = selector(points, bpoints[:, 1])
    b_cell_cent = selector(centroid, bcells)
    b_face_cent = (b_p0 + b_p1) / 2.
    bnx, bny = get_face_norm(b_cell_cent, b_p0, b_p1)
    bdist = b_face_cent - b_cell_cent
    bdist = bdist / (bdist ** 2).sum(dim=-1, keepdim=True)
    Ac = Ac - reducer(bdist[:, 0] * bnx + bdist[:, 1] * bny, bcells, nc)
    return dict(Af=Af, Ac=Ac, b=b, rho=rho, centroid=centroid, area=area)
