# Portions of this file reproduce EASIER (jack-pan-ai/EASIER),
# Copyright (c) Microsoft Corporation, MIT License; the licence text is in
# easier_amd/THIRD_PARTY_NOTICES.txt.
"""Shallow-water geometry and RK4 step restated with torch CPU operators.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

geometry(): tutorial/shallow_water_equation/assemble_shallow_water.py:
  ShallowWaterMeshComponentsCollector.forward (:248-263) then
  ShallowWaterInitializer.forward (:371-439), with the Gaussian's centre,
  width and amplitude as parameters (the tutorial: centre (0, 0), 100, 0.1).
step(): swe_main.py:86-141 with the gravity g as a parameter
  (the tutorial hard-codes g = 1: 0.5*h_f*h_f and 0.5*h_b**2).
"""
import torch

from .ops import reducer, selector


def get_alpha(sc, dc, p0, p1):
    """assemble_shallow_water.py:340-357."""
    x1, y1 = sc[:, 0], sc[:, 1]
    x2, y2 = dc[:, 0], dc[:, 1]
    x3, y3 = p0[:, 0], p0[:, 1]
    x4, y4 = p1[:, 0], p1[:, 1]
    y21 = y2 - y1
    y43 = y4 - y3
    y31 = y3 - y1
    x31 = x3 - x1
    x21 = x2 - x1
    x43 = x4 - x3
    return (x31 * y43 - y31 * x43) / (x21 * y43 - y21 * x43)


def get_face_norm(p0, p1, p2):
    """assemble_shallow_water.py:359-369."""
    a1, a2 = p0[:, 0], p0[:, 1]
    b1, b2 = p1[:, 0], p1[:, 1]
    c1, c2 = p2[:, 0], p2[:, 1]
    s = torch.sign((b1 - c1) * (a2 - c2) - (b2 - c2) * (a1 - c1))
    return s * (b2 - c2), -s * (b1 - c1)


def geometry(mesh, centre=(0.0, 0.0), width=100.0, amp=0.1):
    src = torch.as_tensor(mesh['src'])
    dst = torch.as_tensor(mesh['dst'])
    cells = torch.as_tensor(mesh['cells'])
    points = torch.as_tensor(mesh['points'], dtype=torch.float64)
    bcells = torch.as_tensor(mesh['bcells'])
    bpoints = torch.as_tensor(mesh['bpoints'])
    src_p = selector(cells, src)
    dst_p = selector(cells, dst)
    sp = [selector(points, src_p[:, i]) for i in range(3)]
    dp = [selector(points, dst_p[:, i]) for i in range(3)]
    src_cent = (sp[0] + sp[1] + sp[2]) / 3.
    dst_cent = (dp[0] + dp[1] + dp[2]) / 3.
    n01 = get_face_norm(sp[2], sp[0], sp[1])
    n12 = get_face_norm(sp[0], sp[1], sp[2])
    n20 = get_face_norm(sp[1], sp[2], sp[0])
    n01_ = get_face_norm(dst_cent, sp[0], sp[1])
    n12_ = get_face_norm(dst_cent, sp[1], sp[2])
    n20_ = get_face_norm(dst_cent, sp[2], sp[0])
    ne = src.shape[0]
    cond = (n01[0] * n01_[0] + n01[1] * n01_[1]) < 0
    sx = torch.where(cond, n01[0], 0.)
    sy = torch.where(cond, n01[1], 0.)
    alpha = torch.where(cond, get_alpha(src_cent, dst_cent, sp[0], sp[1]), 0.)
    cond = (n12[0] * n12_[0] + n12[1] * n12_[1]) < 0
    sx = torch.where(cond, n12[0], sx)
    sy = torch.where(cond, n12[1], sy)
    alpha = torch.where(cond, get_alpha(src_cent, dst_cent, sp[1], sp[2]), alpha)
    cond = (n20[0] * n20_[0] + n20[1] * n20_[1]) < 0
    sx = torch.where(cond, n20[0], sx)
    sy = torch.where(cond, n20[1], sy)
    alpha = torch.where(cond, get_alpha(src_cent, dst_cent, sp[2], sp[0]), alpha)
    assert sx.shape[0] == ne
    p0 = selector(points, cells[:, 0])
    p1 = selector(points, cells[:, 1])
    p2 = selector(points, cells[:, 2])
    x0, y0 = p0[:, 0], p0[:, 1]
    x1, y1 = p1[:, 0], p1[:, 1]
    x2, y2 = p2[:, 0], p2[:, 1]
    area = 0.5 * torch.abs(x0 * (y1 - y2) + x1 * (y2 - y0) + x2 * (y0 - y1))
    centroid = (p0 + p1 + p2) / 3.
    x = centroid[:, 0].clone()
    y = centroid[:, 1].clone()
    h = 1 + amp * torch.exp(
        -width * ((x - centre[0]) ** 2 + (y - centre[1]) ** 2))
    b_p0 = selector(points, bpoints[:, 0])
    b_p1 = selector(points, bpoints[:, 1])
    b_cell_cent = selector(centroid, bcells)
    bnx, bny = get_face_norm(b_cell_cent, b_p0, b_p1)
    return dict(x=x, y=y, area=area, sx=sx, sy=sy, bsx=-bnx, bsy=-bny, h=h,
                alpha=alpha)


class SWE:
    """swe_main.py:13-141: state h, uh, vh per cell; one __call__ = one RK4
    step of size dt."""

    def __init__(self, mesh, geo, dt=0.005, g=1.0, dtype=torch.float64):
        self.src = torch.as_tensor(mesh['src'])
        self.dst = torch.as_tensor(mesh['dst'])
        self.bcells = torch.as_tensor(mesh['bcells'])
        self.nc = mesh['cells'].shape[0]
        self.dt, self.g = dt, g
        cast = lambda t: torch.as_tensor(t).to(dtype).clone()
        self.area, self.sx, self.sy = cast(geo['area']), cast(geo['sx']), cast(geo['sy'])
        self.bsx, self.bsy, self.alpha = cast(geo['bsx']), cast(geo['bsy']), cast(geo['alpha'])
        self.h = cast(geo['h'])
        self.uh = torch.zeros(self.nc, dtype=dtype)
        self.vh = torch.zeros(self.nc, dtype=dtype)

    def scatter(self, t):
        return reducer(t, self.dst, self.nc)

    def scatter_b(self, t):
        return reducer(t, self.bcells, self.nc)

    def face_reconstruct(self, phi):
        """swe_main.py:86-88."""
        return (1 - self.alpha) * selector(phi, self.src) + \
            self.alpha * selector(phi, self.dst)

    def delta(self, h, uh, vh):
        """swe_main.py:90-119 (0.5 -> 0.5*g)."""
        h_f = self.face_reconstruct(h)
        uh_f = self.face_reconstruct(uh)
        vh_f = self.face_reconstruct(vh)
        u_f = uh_f / h_f
        v_f = vh_f / h_f
        hg = 0.5 * self.g
        h_f_square = hg * h_f * h_f
        uh_f_times_sx = uh_f * self.sx
        vh_f_times_sy = vh_f * self.sy
        gather_b_h_square = hg * selector(h, self.bcells) ** 2
        delta_h = - (self.scatter(uh_f_times_sx + vh_f_times_sy)) / self.area
        delta_uh = - (
            self.scatter((u_f * uh_f + h_f_square) * self.sx +
                         u_f * vh_f_times_sy) +
            self.scatter_b(gather_b_h_square * self.bsx)) / self.area
        delta_vh = - (
            self.scatter(v_f * uh_f_times_sx +
                         (v_f * vh_f + h_f_square) * self.sy) +
            self.scatter_b(gather_b_h_square * self.bsy)) / self.area
        return delta_h, delta_uh, delta_vh

    def __call__(self):
        """swe_main.py:121-141."""
        dt = self.dt
        d1 = self.delta(self.h, self.uh, self.vh)
        d2 = self.delta(self.h + 0.5 * dt * d1[0], self.uh + 0.5 * dt * d1[1],
                        self.vh + 0.5 * dt * d1[2])
        d3 = self.delta(self.h + 0.5 * dt * d2[0], self.uh + 0.5 * dt * d2[1],
                        self.vh + 0.5 * dt * d2[2])
        d4 = self.delta(self.h + dt * d3[0], self.uh + dt * d3[1],
                        self.vh + dt * d3[2])
        self.h[:] += dt / 6 * (d1[0] + d2[0] + d3[0] + d4[0])
        self.uh[:] += dt / 6 * (d1[1] + d2[1] + d3[1] + d4[1])
        self.vh[:] += dt / 6 * (d1[2] + d2[2] + d3[2] + d4[2])
