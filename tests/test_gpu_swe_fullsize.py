"""The shallow-water step at the benchmark's size, 2-triangle mesh 2048 x 2048
(8,388,608 cells, 25,157,632 directed edges), Gaussian centre drawn from
seed 0 as in the benchmark, float64: one RK4 step
bit-identical to the oracle's (same operations in the same order), total
mass conserved over 20 steps (the directed edges of a face carry opposite
fluxes and the walls carry none), and a lake at rest that stays at rest
(a triangle's outward face normals sum to zero, the wall term cancels the
boundary faces)."""
import math

import pytest
import torch

import easier_amd as esr
from easier_amd.workloads import mesh as wmesh
from easier_amd.workloads import swe as wswe
from oracle import swe as oswe

pytestmark = pytest.mark.gpu

N = 2048
G = 9.81


@pytest.fixture(scope='module')
def full():
    mesh = wmesh.tri2(N)
    geo = oswe.geometry(mesh, centre=wswe.gaussian_centre(0), width=200.0, amp=0.1)
    dt = wswe.cfl_dt(mesh, geo['h'], G)            # CFL 0.4, as the benchmark
    return mesh, geo, dt


def _mass(h, area):
    return math.fsum((h * area).tolist())


def test_full_size_step_bit_identical_and_mass_conserved(gpu, full):
    mesh, geo, dt = full
    print(f'swe full size {N}x{N} step vs oracle', flush=True)
    eqn = wswe.ShallowWaterEquation(mesh, geo, dt=dt, g=G, dtype=torch.float64)
    esr.compile([eqn], 'hip')
    ref = oswe.SWE(mesh, geo, dt=dt, g=G, dtype=torch.float64)
    area = torch.as_tensor(geo['area'], dtype=torch.float64)
    m0 = _mass(torch.as_tensor(geo['h'], dtype=torch.float64), area)
    eqn()
    ref()
    for k in ('h', 'uh', 'vh'):
        assert torch.equal(getattr(eqn, k).collect(), getattr(ref, k)), k
    del ref
    for _ in range(19):
        eqn()
    h = eqn.h.collect()
    assert bool(torch.isfinite(h).all())
    m20 = _mass(h, area)
    assert abs(m20 - m0) <= 1e-12 * abs(m0), (m0, m20)


@pytest.mark.parametrize('reorder', ['rcm', 'blocks'])
def test_full_size_frozen_step_bit_identical(gpu, full, reorder):
    """compile(..., freeze=True) reads the geometry through dictionaries
    (the uniform mesh's alpha, sx, sy take few values; area one): the same
    bits as the oracle after two steps."""
    mesh, geo, dt = full
    print(f'swe full size {N}x{N} frozen steps vs oracle, reorder={reorder}', flush=True)
    eqn = wswe.ShallowWaterEquation(mesh, geo, dt=dt, g=G, dtype=torch.float64)
    esr.compile([eqn], 'hip', freeze=True, reorder=reorder)
    assert any(k.luts for k in eqn.easier_engine.kernels)
    ref = oswe.SWE(mesh, geo, dt=dt, g=G, dtype=torch.float64)
    for _ in range(2):
        eqn()
        ref()
    for k in ('h', 'uh', 'vh'):
        assert torch.equal(getattr(eqn, k).collect(), getattr(ref, k)), k


def test_full_size_lake_at_rest(gpu, full):
    mesh, geo, dt = full
    print(f'swe full size {N}x{N} lake at rest', flush=True)
    flat = dict(geo, h=torch.ones_like(torch.as_tensor(geo['h'])))
    eqn = wswe.ShallowWaterEquation(mesh, flat, dt=dt, g=G, dtype=torch.float64)
    esr.compile([eqn], 'hip')
    for _ in range(10):
        eqn()
    h, uh, vh = eqn.h.collect(), eqn.uh.collect(), eqn.vh.collect()
    assert float((h - 1.0).abs().max()) <= 1e-12
    assert float(uh.abs().max()) <= 1e-10 and float(vh.abs().max()) <= 1e-10


def test_full_size_float32_step_matches_oracle(gpu, full):
    mesh, geo, dt = full
    print(f'swe full size {N}x{N} float32 step vs oracle', flush=True)
    eqn = wswe.ShallowWaterEquation(mesh, geo, dt=dt, g=G, dtype=torch.float32)
    esr.compile([eqn], 'hip')
    ref = oswe.SWE(mesh, geo, dt=dt, g=G, dtype=torch.float32)
    for _ in range(2):
        eqn()
        ref()
    for k in ('h', 'uh', 'vh'):
        got, want = getattr(eqn, k).collect(), getattr(ref, k)
        scale = float(want.abs().max())
        assert float((got - want).abs().max()) <= 1e-5 * scale, k
