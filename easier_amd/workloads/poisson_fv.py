"""The reference's Poisson tutorial problem: a cell-centred two-point-flux
finite-volume operator on a triangle mesh (the problem its integration test
solves with CG and GMRES, tests/test_numeric.py:14-68).

FVPoissonAssembler computes, as an esr program run by the hip backend, the
operator of tutorial/poisson/assemble_poisson.py:160-230. Per directed edge
e = (s, t) of the cell graph, with centroids c_s, c_t and the face the two
cells share:

    Af[e] = (c_t - c_s) . n / |c_t - c_s|^2,

n the face's normal pointing out of s, scaled by the face length. Per cell,
Ac = -(sum of Af over the cell's outgoing edges) - (sum over its wall faces
of (m - c) . n / |m - c|^2, m the face midpoint, n the outward normal), and
b = rho * area with rho = exp(-200 |c - (0.5, 0.5)|^2).

The shared face is found geometrically, as in the tutorial: of the source
triangle's three edges it is the one whose line separates the source
triangle's own third vertex from the destination centroid (side tests by
cross products); where a distorted mesh puts the destination centroid beyond
two edge lines, the later edge in the order (v0 v1, v1 v2, v2 v0) is taken,
as the tutorial's sequence of torch.where calls does. FVPoisson is tutorial/poisson/poisson_main.py:16-53: x, b,
Ac, Af and A = Linsys(Ac, Af, Selector(dst), Reducer(src)).
"""
import numpy as np
import torch

import easier_amd as esr
from ..numeric import Linsys


def _cross(ax, ay, bx, by):
    return ax * by - ay * bx


class FVPoissonAssembler(esr.Module):
    def __init__(self, mesh, centre=(0.5, 0.5)):
        super().__init__()
        cells = np.asarray(mesh['cells'])
        src, dst = np.asarray(mesh['src']), np.asarray(mesh['dst'])
        bcells, bpoints = np.asarray(mesh['bcells']), np.asarray(mesh['bpoints'])
        nc, ne = cells.shape[0], src.shape[0]
        # vertex k of the edge's source / destination triangle, of each cell,
        # and of each wall face (what the tutorial's collector program writes)
        pick = lambda a: torch.from_numpy(np.ascontiguousarray(a))
        self.px = esr.Tensor(torch.from_numpy(mesh['points'][:, 0].copy()), mode='partition')
        self.py = esr.Tensor(torch.from_numpy(mesh['points'][:, 1].copy()), mode='partition')
        self.src_v = torch.nn.ModuleList(
            [esr.Selector(pick(cells[src, k])) for k in range(3)])
        self.dst_v = torch.nn.ModuleList(
            [esr.Selector(pick(cells[dst, k])) for k in range(3)])
        self.cell_v = torch.nn.ModuleList(
            [esr.Selector(pick(cells[:, k])) for k in range(3)])
        self.wall_v = torch.nn.ModuleList(
            [esr.Selector(pick(bpoints[:, k])) for k in range(2)])
        self.wall_cell = esr.Selector(pick(bcells))
        self.out_edges = esr.Reducer(pick(src), nc)
        self.wall_faces = esr.Reducer(pick(bcells), nc)
        self.cx0, self.cy0 = float(centre[0]), float(centre[1])
        z = lambda k: esr.Tensor(esr.zeros([k], dtype=torch.double), mode='partition')
        self.Af, self.Ac, self.b, self.rho = z(ne), z(nc), z(nc), z(nc)
        self.cx, self.cy, self.area = z(nc), z(nc), z(nc)

    def _xy(self, sel):
        return sel(self.px), sel(self.py)

    def forward(self):
        # cells: centroid and area
        (x0, y0), (x1, y1), (x2, y2) = [self._xy(s) for s in self.cell_v]
        cx = (x0 + x1 + x2) / 3.
        cy = (y0 + y1 + y2) / 3.
        self.cx[:] = cx
        self.cy[:] = cy
        self.area[:] = 0.5 * torch.abs(_cross(x1 - x0, y1 - y0, x2 - x0, y2 - y0))
        self.rho[:] = torch.exp(-200. * ((cx - self.cx0) ** 2 + (cy - self.cy0) ** 2))
        self.b[:] = self.rho * self.area

        # edges: the face between the two cells and the two-point flux
        sv = [self._xy(s) for s in self.src_v]
        tv = [self._xy(s) for s in self.dst_v]
        scx = (sv[0][0] + sv[1][0] + sv[2][0]) / 3.
        scy = (sv[0][1] + sv[1][1] + sv[2][1]) / 3.
        tcx = (tv[0][0] + tv[1][0] + tv[2][0]) / 3.
        tcy = (tv[0][1] + tv[1][1] + tv[2][1]) / 3.
        nx = torch.zeros_like(scx)
        ny = torch.zeros_like(scx)
        for a, b, o in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
            (ax, ay), (bx, by), (ox, oy) = sv[a], sv[b], sv[o]
            ex, ey = bx - ax, by - ay
            side_o = _cross(ex, ey, ox - ax, oy - ay)
            side_t = _cross(ex, ey, tcx - ax, tcy - ay)
            shared = side_o * side_t < 0
            s = torch.sign(side_o)                   # n = s (ey, -ex) points away from o
            nx = torch.where(shared, s * ey, nx)
            ny = torch.where(shared, -s * ex, ny)
        dx, dy = tcx - scx, tcy - scy
        self.Af[:] = (dx * nx + dy * ny) / (dx * dx + dy * dy)

        # wall faces
        (ax, ay), (bx, by) = [self._xy(s) for s in self.wall_v]
        wcx, wcy = self.wall_cell(cx), self.wall_cell(cy)
        ex, ey = bx - ax, by - ay
        s = torch.sign(_cross(ex, ey, wcx - ax, wcy - ay))    # n = s (ey, -ex) points away from the cell
        mx, my = 0.5 * (ax + bx) - wcx, 0.5 * (ay + by) - wcy
        wall = (mx * s * ey - my * s * ex) / (mx * mx + my * my)
        self.Ac[:] = -self.out_edges(self.Af) - self.wall_faces(wall)


def assemble(mesh, centre=(0.5, 0.5)):
    """Run the assembler on the GPU; host tensors in mesh order."""
    asm = FVPoissonAssembler(mesh, centre)
    [asm] = esr.compile([asm], 'hip')
    asm()
    return {k: getattr(asm, k).collect() for k in
            ('Af', 'Ac', 'b', 'rho', 'cx', 'cy', 'area')}


class FVPoisson(esr.Module):
    """poisson_main.py:16-53 on in-memory arrays: Linsys(Ac, Af,
    Selector(dst), Reducer(src))."""

    def __init__(self, mesh, system, x=None):
        super().__init__()
        nc = np.asarray(mesh['cells']).shape[0]
        self.reducer = esr.Reducer(torch.from_numpy(np.asarray(mesh['src'])), nc)
        self.selector = esr.Selector(torch.from_numpy(np.asarray(mesh['dst'])))
        self.x = esr.Tensor(esr.zeros((nc,), dtype=torch.double), mode='partition') \
            if x is None else x
        as64 = lambda t: torch.as_tensor(t, dtype=torch.float64).clone()
        self.b = esr.Tensor(as64(system['b']), mode='partition')
        self.Ac = esr.Tensor(as64(system['Ac']), mode='partition')
        self.Af = esr.Tensor(as64(system['Af']), mode='partition')
        self.A = Linsys(self.Ac, self.Af, self.selector, self.reducer)
