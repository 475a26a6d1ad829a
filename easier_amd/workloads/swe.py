# Portions of this file reproduce EASIER (jack-pan-ai/EASIER),
# Copyright (c) Microsoft Corporation, MIT License; the licence text is in
# easier_amd/THIRD_PARTY_NOTICES.txt.
"""The shallow-water tutorial programs on in-memory meshes.

ShallowWaterMeshComponentsCollector and ShallowWaterInitializer follow
tutorial/shallow_water_equation/assemble_shallow_water.py:196-439 (the
Gaussian's centre, width and amplitude are arguments; the tutorial uses
(0, 0), 100, 0.1); ShallowWaterEquation follows swe_main.py:13-141 with the
gravity g as an argument (the tutorial's fluxes are those of g = 1: 0.5*h*h).
"""
import math

import torch

import easier_amd as esr


class ShallowWaterMeshComponentsCollector(esr.Module):
    def __init__(self, mesh):
        super().__init__()
        self.cells = esr.Tensor(torch.from_numpy(mesh['cells']), mode='partition')
        self.bpoints = esr.Tensor(torch.from_numpy(mesh['bpoints']),
                                  mode='partition')
        self.selector_src = esr.Selector(torch.from_numpy(mesh['src']))
        self.selector_dst = esr.Selector(torch.from_numpy(mesh['dst']))
        nc = mesh['cells'].shape[0]
        ne = mesh['src'].shape[0]
        nbc = mesh['bpoints'].shape[0]
        self.src_p = torch.nn.ParameterList([
            esr.Tensor(esr.zeros([ne], dtype=torch.long), mode='partition')
            for i in range(3)])
        self.dst_p = torch.nn.ParameterList([
            esr.Tensor(esr.zeros([ne], dtype=torch.long), mode='partition')
            for i in range(3)])
        self.cells_p = torch.nn.ParameterList([
            esr.Tensor(esr.zeros([nc], dtype=torch.long), mode='partition')
            for i in range(3)])
        self.bp = torch.nn.ParameterList([
            esr.Tensor(esr.zeros([nbc], dtype=torch.long), mode='partition')
            for i in range(2)])

    def forward(self):
        src_p = self.selector_src(self.cells)
        dst_p = self.selector_dst(self.cells)
        for i in range(3):
            self.src_p[i].copy_(src_p[:, i])
            self.dst_p[i].copy_(dst_p[:, i])
            self.cells_p[i].copy_(self.cells[:, i])
        for i in range(2):
            self.bp[i].copy_(self.bpoints[:, i])


class ShallowWaterInitializer(esr.Module):
    def __init__(self, mesh, components, centre=(0.0, 0.0), width=100.0,
                 amp=0.1):
        super().__init__()
        self.points = esr.Tensor(torch.from_numpy(mesh['points']),
                                 mode='partition')
        nc = mesh['cells'].shape[0]
        ne = mesh['src'].shape[0]
        nbc = mesh['bcells'].shape[0]
        self.selector_src_p = torch.nn.ModuleList(
            [esr.Selector(components['src_p'][i]) for i in range(3)])
        self.selector_dst_p = torch.nn.ModuleList(
            [esr.Selector(components['dst_p'][i]) for i in range(3)])
        self.selector_cells_p = torch.nn.ModuleList(
            [esr.Selector(components['cells_p'][i]) for i in range(3)])
        self.bselector = esr.Selector(torch.from_numpy(mesh['bcells']))
        self.selector_bp = torch.nn.ModuleList(
            [esr.Selector(components['bp'][i]) for i in range(2)])
        self.centre = (float(centre[0]), float(centre[1]))
        self.width, self.amp = float(width), float(amp)
        dz = lambda k: esr.Tensor(esr.zeros([k], dtype=torch.double),
                                  mode='partition')
        self.x, self.y, self.area = dz(nc), dz(nc), dz(nc)
        self.sx, self.sy = dz(ne), dz(ne)
        self.bsx, self.bsy = dz(nbc), dz(nbc)
        self.h = dz(nc)
        self.alpha = dz(ne)

    def get_alpha(self, sc, dc, p0, p1):
        x1 = sc[:, 0]
        y1 = sc[:, 1]
        x2 = dc[:, 0]
        y2 = dc[:, 1]
        x3 = p0[:, 0]
        y3 = p0[:, 1]
        x4 = p1[:, 0]
        y4 = p1[:, 1]
        y21 = y2 - y1
        y43 = y4 - y3
        y31 = y3 - y1
        x31 = x3 - x1
        x21 = x2 - x1
        x43 = x4 - x3
        return (x31 * y43 - y31 * x43) / (x21 * y43 - y21 * x43)

    def get_face_norm(self, p0, p1, p2):
        a1 = p0[:, 0]
        a2 = p0[:, 1]
        b1 = p1[:, 0]
        b2 = p1[:, 1]
        c1 = p2[:, 0]
        c2 = p2[:, 1]
        s = torch.sign((b1 - c1) * (a2 - c2) - (b2 - c2) * (a1 - c1))
        return s * (b2 - c2), -s * (b1 - c1)

    def forward(self):
        src_p0 = self.selector_src_p[0](self.points)
        src_p1 = self.selector_src_p[1](self.points)
        src_p2 = self.selector_src_p[2](self.points)
        dst_p0 = self.selector_dst_p[0](self.points)
        dst_p1 = self.selector_dst_p[1](self.points)
        dst_p2 = self.selector_dst_p[2](self.points)
        src_cent = (src_p0 + src_p1 + src_p2) / 3.
        dst_cent = (dst_p0 + dst_p1 + dst_p2) / 3.
        norm01_x, norm01_y = self.get_face_norm(src_p2, src_p0, src_p1)
        norm12_x, norm12_y = self.get_face_norm(src_p0, src_p1, src_p2)
        norm20_x, norm20_y = self.get_face_norm(src_p1, src_p2, src_p0)
        norm01_x_, norm01_y_ = self.get_face_norm(dst_cent, src_p0, src_p1)
        norm12_x_, norm12_y_ = self.get_face_norm(dst_cent, src_p1, src_p2)
        norm20_x_, norm20_y_ = self.get_face_norm(dst_cent, src_p2, src_p0)

        condition = (norm01_x * norm01_x_ + norm01_y * norm01_y_) < 0
        self.sx[:] = torch.where(condition, norm01_x, 0.)
        self.sy[:] = torch.where(condition, norm01_y, 0.)
        alpha = self.get_alpha(src_cent, dst_cent, src_p0, src_p1)
        self.alpha[:] = torch.where(condition, alpha, 0.)

        condition = (norm12_x * norm12_x_ + norm12_y * norm12_y_) < 0
        self.sx[:] = torch.where(condition, norm12_x, self.sx)
        self.sy[:] = torch.where(condition, norm12_y, self.sy)
        alpha = self.get_alpha(src_cent, dst_cent, src_p1, src_p2)
        self.alpha[:] = torch.where(condition, alpha, self.alpha)

        condition = (norm20_x * norm20_x_ + norm20_y * norm20_y_) < 0
        self.sx[:] = torch.where(condition, norm20_x, self.sx)
        self.sy[:] = torch.where(condition, norm20_y, self.sy)
        alpha = self.get_alpha(src_cent, dst_cent, src_p2, src_p0)
        self.alpha[:] = torch.where(condition, alpha, self.alpha)

        p0 = self.selector_cells_p[0](self.points)
        x0 = p0[:, 0]
        y0 = p0[:, 1]
        p1 = self.selector_cells_p[1](self.points)
        x1 = p1[:, 0]
        y1 = p1[:, 1]
        p2 = self.selector_cells_p[2](self.points)
        x2 = p2[:, 0]
        y2 = p2[:, 1]
        self.area[:] = 0.5 * torch.abs(
            x0 * (y1 - y2) + x1 * (y2 - y0) + x2 * (y0 - y1))
        centroid = (p0 + p1 + p2) / 3.
        self.x[:] = centroid[:, 0]
        self.y[:] = centroid[:, 1]
        self.h[:] = 1 + self.amp * torch.exp(
            -self.width * ((self.x - self.centre[0])**2 +
                           (self.y - self.centre[1])**2))
        b_p0 = self.selector_bp[0](self.points)
        b_p1 = self.selector_bp[1](self.points)
        b_cell_cent = self.bselector(centroid)
        bnorm_x, bnorm_y = self.get_face_norm(b_cell_cent, b_p0, b_p1)
        self.bsx[:] = -bnorm_x
        self.bsy[:] = -bnorm_y


def assemble_geometry(mesh, centre=(0.0, 0.0), width=100.0, amp=0.1):
    """Run both assembly programs on the GPU; returns host tensors in mesh
    order (assemble_shallow_water.py:442-478, without the HDF5 files)."""
    comp = ShallowWaterMeshComponentsCollector(mesh)
    [comp] = esr.compile([comp], 'hip')
    comp()
    components = {
        'src_p': [comp.src_p[i].collect() for i in range(3)],
        'dst_p': [comp.dst_p[i].collect() for i in range(3)],
        'cells_p': [comp.cells_p[i].collect() for i in range(3)],
        'bp': [comp.bp[i].collect() for i in range(2)],
    }
    init = ShallowWaterInitializer(mesh, components, centre, width, amp)
    [init] = esr.compile([init], 'hip')
    init()
    return {k: getattr(init, k).collect() for k in
            ('x', 'y', 'area', 'sx', 'sy', 'bsx', 'bsy', 'h', 'alpha')}


class ShallowWaterEquation(esr.Module):
    """swe_main.py:13-141 (g generalised; g = 1 is the tutorial)."""

    def __init__(self, mesh, geo, dt=0.005, g=1.0, dtype=torch.float64):
        super().__init__()
        self.dt = float(dt)
        self.g = float(g)
        src = torch.from_numpy(mesh['src'])
        dst = torch.from_numpy(mesh['dst'])
        bcells = torch.from_numpy(mesh['bcells'])
        self.nc = mesh['cells'].shape[0]
        self.scatter = esr.Reducer(dst, self.nc)
        self.gather_src = esr.Selector(src)
        self.gather_dst = esr.Selector(dst)
        self.scatter_b = esr.Reducer(bcells, self.nc)
        self.gather_b = esr.Selector(bcells)
        T = lambda k: esr.Tensor(torch.as_tensor(geo[k]).to(dtype),
                                 mode='partition')
        self.x, self.y, self.area = T('x'), T('y'), T('area')
        self.sx, self.sy, self.bsx, self.bsy = T('sx'), T('sy'), T('bsx'), T('bsy')
        self.h, self.alpha = T('h'), T('alpha')
        self.uh = esr.Tensor(esr.zeros((self.nc,), dtype=dtype), mode='partition')
        self.vh = esr.Tensor(esr.zeros((self.nc,), dtype=dtype), mode='partition')

    def face_reconstruct(self, phi):
        return (1 - self.alpha) * self.gather_src(phi) + \
            self.alpha * self.gather_dst(phi)

    def delta(self, h, uh, vh):
        h_f = self.face_reconstruct(h)
        uh_f = self.face_reconstruct(uh)
        vh_f = self.face_reconstruct(vh)
        u_f = uh_f / h_f
        v_f = vh_f / h_f
        h_f_square = 0.5 * self.g * h_f * h_f
        uh_f_times_sx = uh_f * self.sx
        vh_f_times_sy = vh_f * self.sy
        gather_b_h_square = 0.5 * self.g * self.gather_b(h)**2
        delta_h = - (
            self.scatter(uh_f_times_sx + vh_f_times_sy)
        ) / self.area
        delta_uh = - (
            self.scatter((u_f * uh_f + h_f_square) * self.sx +
                         u_f * vh_f_times_sy) +
            self.scatter_b(gather_b_h_square * self.bsx)
        ) / self.area
        delta_vh = - (
            self.scatter(v_f * uh_f_times_sx +
                         (v_f * vh_f + h_f_square) * self.sy) +
            self.scatter_b(gather_b_h_square * self.bsy)
        ) / self.area
        return delta_h, delta_uh, delta_vh

    def forward(self):
        delta_h1, delta_uh1, delta_vh1 = self.delta(self.h, self.uh, self.vh)
        delta_h2, delta_uh2, delta_vh2 = self.delta(
            self.h + 0.5 * self.dt * delta_h1,
            self.uh + 0.5 * self.dt * delta_uh1,
            self.vh + 0.5 * self.dt * delta_vh1,)
        delta_h3, delta_uh3, delta_vh3 = self.delta(
            self.h + 0.5 * self.dt * delta_h2,
            self.uh + 0.5 * self.dt * delta_uh2,
            self.vh + 0.5 * self.dt * delta_vh2,)
        delta_h4, delta_uh4, delta_vh4 = self.delta(
            self.h + self.dt * delta_h3,
            self.uh + self.dt * delta_uh3,
            self.vh + self.dt * delta_vh3,)
        self.h[:] += self.dt / 6 * (
            delta_h1 + delta_h2 + delta_h3 + delta_h4)
        self.uh[:] += self.dt / 6 * (
            delta_uh1 + delta_uh2 + delta_uh3 + delta_uh4)
        self.vh[:] += self.dt / 6 * (
            delta_vh1 + delta_vh2 + delta_vh3 + delta_vh4)


def gaussian_centre(seed):
    """Centre x0 of the initial Gaussian, uniform in [0.3, 0.7]^2 from numpy's
    PCG64 seeded with `seed`."""
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(seed))
    return tuple(float(v) for v in rng.uniform(0.3, 0.7, 2))


def cfl_dt(mesh, h0, g, cfl=0.4):
    """dt = cfl * L_min / (max|u| + sqrt(g * h_max)) at rest (u = 0), with
    L_min the shortest mesh edge."""
    pts = mesh['points']
    cells = mesh['cells']
    lmin = math.inf
    for a, b in ((0, 1), (1, 2), (2, 0)):
        d = pts[cells[:, a]] - pts[cells[:, b]]
        lmin = min(lmin, float(((d ** 2).sum(1) ** 0.5).min()))
    return cfl * lmin / math.sqrt(g * float(torch.as_tensor(h0).max()))
