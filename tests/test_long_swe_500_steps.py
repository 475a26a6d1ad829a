"""The benchmark's full 500-step shallow-water horizon at a size the host
oracle runs in seconds: 2-triangle mesh 128 x 128 (32,768 cells, 97,792
directed edges), Gaussian centre drawn from seed 0 (x0 ~ U[0.3, 0.7]^2),
g = 9.81, CFL 0.4, reflective walls. The compiled (fused) step runs 500 times
on the GPU; float64 must stay bit-identical to the oracle's torch CPU steps
after every 100 steps, float32 within 1e-5 of the oracle's float32 run
relative to the field's magnitude. These run last among the GPU tests (the
longest)."""
import pytest
import torch

import easier_amd as esr
from easier_amd.workloads import mesh as wmesh
from easier_amd.workloads import swe as wswe
from oracle import swe as oswe

pytestmark = pytest.mark.gpu

N, STEPS, G = 128, 500, 9.81


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_500_steps_against_the_oracle(gpu, dtype):
    print(f'swe {N}x{N} {STEPS} steps {dtype}', flush=True)
    mesh = wmesh.tri2(N)
    geo = oswe.geometry(mesh, centre=wswe.gaussian_centre(0), width=200.0, amp=0.1)
    dt = wswe.cfl_dt(mesh, geo['h'], G)
    eqn = wswe.ShallowWaterEquation(mesh, geo, dt=dt, g=G, dtype=dtype)
    esr.compile([eqn], 'hip')
    ref = oswe.SWE(mesh, geo, dt=dt, g=G, dtype=dtype)
    for step in range(1, STEPS + 1):
        eqn()
        ref()
        if step % 100:
            continue
        for k in ('h', 'uh', 'vh'):
            got, want = getattr(eqn, k).collect(), getattr(ref, k)
            if dtype == torch.float64:
                assert torch.equal(got, want), (step, k)
            else:
                scale = float(want.abs().max())
                err = float((got - want).abs().max())
                assert err <= 1e-5 * scale, (step, k, err, scale)
        print(f'  step {step} ok', flush=True)
    assert float(ref.uh.abs().max()) > 1e-3            # the wave has moved
