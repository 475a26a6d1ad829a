"""End-to-end parity of compiled programs (esr.compile(..., 'hip')) against
the oracle: Linsys/CG/GMRES, the shallow-water assembly programs and RK4
steps, Reducer `out=` semantics and Tensor.collect() ordering."""
import numpy as np
import pytest
import torch

import easier_amd as esr
from easier_amd.numeric import Linsys
from easier_amd.numeric.preconditioner import Jacobi
from easier_amd.numeric.solver import CG, GMRES
from easier_amd.workloads import mesh as wmesh
from easier_amd.workloads import swe as wswe
from oracle import ops as oops
from oracle import solvers as osol
from oracle import swe as oswe

pytestmark = pytest.mark.gpu


def _laplacian_edges(m):
    rows, cols = [], []
    for j in range(m):
        for i in range(m):
            for di, dj in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                a, b = i + di, j + dj
                if 0 <= a < m and 0 <= b < m:
                    rows.append(j * m + i)
                    cols.append(b * m + a)
    return torch.tensor(rows), torch.tensor(cols)


class Poisson5(esr.Module):
    """5-point Laplacian + diagonal shift as a Linsys (tutorial/poisson
    layout: Reducer(rows), Selector(cols), Af per edge, Ac per vertex)."""

    def __init__(self, m, seed=0):
        super().__init__()
        rows, cols = _laplacian_edges(m)
        n = m * m
        g = torch.Generator().manual_seed(seed)
        self.rows, self.cols = rows, cols
        self.reducer = esr.Reducer(rows, n)
        self.selector = esr.Selector(cols)
        self.x = esr.Tensor(esr.zeros((n,), dtype=torch.double), mode='partition')
        self.b = esr.Tensor(torch.rand(n, generator=g, dtype=torch.double) * 2 - 1,
                            mode='partition')
        self.Ac = esr.Tensor(4.0 + torch.rand(n, generator=g, dtype=torch.double),
                             mode='partition')
        self.Af = esr.Tensor(-torch.ones(rows.shape[0], dtype=torch.double),
                             mode='partition')
        self.A = Linsys(self.Ac, self.Af, self.selector, self.reducer)

    def oracle(self):
        n = self.Ac.easier_data_loader.shape[0]
        return osol.Linsys(self.Ac.easier_data_loader.load_cpu(),
                           self.Af.easier_data_loader.load_cpu(),
                           self.cols, self.rows, n), \
            self.b.easier_data_loader.load_cpu()


@pytest.mark.parametrize('precond', [False, True])
@pytest.mark.parametrize('reorder', ['rcm', 'none'])
def test_cg_matches_oracle(gpu, precond, reorder):
    print(f'cg precond={precond} reorder={reorder}', flush=True)
    eqn = Poisson5(24)
    A_o, b_o = eqn.oracle()
    M = Jacobi(eqn.Ac) if precond else None
    sol = CG(eqn.A, eqn.b, eqn.x, M)
    esr.compile([eqn, sol], 'hip', reorder=reorder)
    info = sol.solve(rtol=1e-10, maxiter=2000)
    x_o = torch.zeros_like(b_o)
    info_o, x_o = osol.cg(A_o, b_o, x_o, M=(lambda r: r / A_o.Av) if precond else None,
                          rtol=1e-10, maxiter=2000)
    assert abs(info['iters'] - info_o['iters']) <= 1
    assert info['residual'] <= 1e-10 * float(torch.norm(b_o)) * 1.0000001
    torch.testing.assert_close(eqn.x.collect(), x_o, rtol=1e-8, atol=1e-10)


@pytest.mark.parametrize('m', [24, 370])
def test_cg_final_state_matches_oracle(gpu, m):
    """The fused CG leaves the module's r, p and rzsum as the reference loop
    does at exit: 24^2 unknowns run the two-kernel iteration (p alternates
    between two buffers and is copied back at exit), 370^2 = 136,900 the
    three-kernel one."""
    print(f'cg final state m={m}', flush=True)
    eqn = Poisson5(m, seed=3)
    A_o, b_o = eqn.oracle()
    sol = CG(eqn.A, eqn.b, eqn.x, Jacobi(eqn.Ac))
    esr.compile([eqn, sol], 'hip')
    info = sol.solve(rtol=1e-10, maxiter=2000)
    fin = {}
    x_o = torch.zeros_like(b_o)
    info_o, x_o = osol.cg(A_o, b_o, x_o, M=lambda r: r / A_o.Av, rtol=1e-10, maxiter=2000,
                          final=fin)
    assert info['iters'] == info_o['iters'], (info, info_o)
    torch.testing.assert_close(eqn.x.collect(), x_o, rtol=1e-8, atol=1e-10)
    for name in ('r', 'p'):
        got, want = getattr(sol, name).collect(), fin[name]
        err = float((got - want).norm() / want.norm())
        assert err < 1e-6, (name, err)
    assert abs(float(sol.rzsum.collect()[0]) - fin['rzsum']) <= 1e-6 * abs(fin['rzsum'])


@pytest.mark.parametrize('m,precond', [(12, False), (12, True), (140, False), (140, True)])
def test_gmres_matches_oracle(gpu, m, precond):
    """144 and 19,600 unknowns, identity and Jacobi preconditioner."""
    print(f'gmres m={m} precond={precond}', flush=True)
    eqn = Poisson5(m, seed=1)
    A_o, b_o = eqn.oracle()
    M = Jacobi(eqn.Ac) if precond else None
    sol = GMRES(eqn.A, eqn.b, eqn.x, M, restart=10)
    esr.compile([eqn, sol], 'hip')
    info = sol.solve(rtol=1e-9, maxiter=200)
    x_o = torch.zeros_like(b_o)
    info_o, x_o = osol.gmres(A_o, b_o, x_o, M=(lambda r: r / A_o.Av) if precond else None,
                             restart=10, rtol=1e-9, maxiter=200)
    assert abs(info['iters'] - info_o['iters']) <= 1
    torch.testing.assert_close(eqn.x.collect(), x_o, rtol=1e-7, atol=1e-9)
    assert info['residual'] < 1e-9 * float(torch.norm(b_o)) * 1.0001


@pytest.mark.parametrize('precond,maxiter', [(True, None), (False, 3)])
def test_gmres_device_loop_matches_host_loop(gpu, precond, maxiter):
    """The restart loop on the device (Givens least squares, several cycles
    per graph) and on the host (debug_iter: the reference's thin-SVD least
    squares every cycle) run the same cycles to the same residual, and a cycle
    cap stops both after the same number of cycles."""
    print(f'gmres device vs host loop precond={precond} maxiter={maxiter}', flush=True)
    out = []
    for debug in (None, 1000):
        eqn = Poisson5(60, seed=2)
        M = Jacobi(eqn.Ac) if precond else None
        sol = GMRES(eqn.A, eqn.b, eqn.x, M, restart=10)
        esr.compile([eqn, sol], 'hip')
        info = sol.solve(rtol=1e-10, maxiter=maxiter, debug_iter=debug)
        out.append((info, eqn.x.collect().clone(), sol.rnorm.collect().clone()))
    (a, xa, ra), (b, xb, rb) = out
    assert a['iters'] == b['iters']
    if maxiter is not None:
        assert a['iters'] == maxiter
    # x differs at rounding level between the two least-squares methods,
    # which shows in the last digits of residuals near 1e-10 ||b||
    assert abs(a['residual'] - b['residual']) <= 1e-6 * b['residual']
    assert float(ra[0]) == a['residual']
    torch.testing.assert_close(xa, xb, rtol=1e-9, atol=1e-12)


def test_reducer_out_and_collect_in_program(gpu):
    class M(esr.Module):
        def __init__(self):
            super().__init__()
            idx = torch.tensor([1, 1, 2, 2, 2])
            self.red = esr.Reducer(idx, 3)
            self.sel = esr.Selector(torch.tensor([2, 0, 1, 1, 0]))
            g = torch.Generator().manual_seed(0)
            self.v = esr.Tensor(torch.randn((3, 2), generator=g, dtype=torch.double),
                                mode='partition')
            self.e = esr.Tensor(torch.randn((5, 2), generator=g, dtype=torch.double),
                                mode='partition')
            self.o = esr.Tensor(torch.full((3, 2), 9.0, dtype=torch.double),
                                mode='partition')

        def forward(self):
            self.red(self.sel(self.v) * self.e, out=self.o)

    m = M()
    v0, e0 = m.v.easier_data_loader.load_cpu(), m.e.easier_data_loader.load_cpu()
    with pytest.raises(RuntimeError):
        m.sel(v0)
    esr.compile([m], 'hip')
    m()
    ref = oops.reducer(oops.selector(v0, torch.tensor([2, 0, 1, 1, 0])) * e0,
                       torch.tensor([1, 1, 2, 2, 2]), 3,
                       out=torch.full((3, 2), 9.0, dtype=torch.double))
    assert torch.equal(m.o.collect(), ref)
    with pytest.raises(esr.EasierJitException):
        esr.compile([m], 'hip')


@pytest.mark.parametrize('gen', ['tri4', 'tri2'])
def test_swe_geometry_programs(gpu, gen):
    mesh = getattr(wmesh, gen)(6)
    geo = wswe.assemble_geometry(mesh, centre=(0.4, 0.6), width=30.0, amp=0.2)
    ref = oswe.geometry(mesh, centre=(0.4, 0.6), width=30.0, amp=0.2)
    for k in ('x', 'y', 'area', 'sx', 'sy', 'bsx', 'bsy', 'alpha'):
        assert torch.equal(geo[k].cpu(), ref[k]), k
    torch.testing.assert_close(geo['h'].cpu(), ref['h'], rtol=1e-15, atol=0)


@pytest.mark.parametrize('fuse', [True, False])
@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
@pytest.mark.parametrize('gen,g', [('tri4', 1.0), ('tri2', 9.81)])
def test_swe_steps_match_oracle(gpu, dtype, gen, g, fuse):
    print(f'swe {gen} g={g} {dtype} fuse={fuse}', flush=True)
    mesh = getattr(wmesh, gen)(10)
    geo = oswe.geometry(mesh, centre=(0.45, 0.55), width=200.0, amp=0.1)
    dt = 0.2 * wswe.cfl_dt(mesh, geo['h'], g)
    eqn = wswe.ShallowWaterEquation(mesh, geo, dt=dt, g=g, dtype=dtype)
    esr.compile([eqn], 'hip', fuse=fuse)
    assert (type(eqn.easier_engine).__name__ == 'FusedEngine') == fuse
    ref = oswe.SWE(mesh, geo, dt=dt, g=g, dtype=dtype)
    for _ in range(20):           # from the 2nd call on: graph replays
        eqn()
        ref()
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    for k in ('h', 'uh', 'vh'):
        got, want = getattr(eqn, k).collect(), getattr(ref, k)
        scale = float(want.abs().max())
        assert float((got - want).abs().max()) <= tol * scale, k
        if dtype == torch.float64:
            # same operations in the same order, no FMA contraction
            assert torch.equal(got, want), k


@pytest.mark.parametrize('reorder', ['rcm', 'blocks'])
@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
@pytest.mark.parametrize('gen', ['tri2', 'tri4', 'tri2_jitter'])
def test_frozen_swe_steps_match_oracle(gpu, dtype, gen, reorder):
    """compile(..., freeze=True): the geometry is read through dictionaries
    (uniform meshes) or as arrays (jittered: too many distinct values), and
    with reorder='blocks' the gathers read per-tile LDS halos; 30 steps
    bit-identical to the oracle in float64, within 1e-5 in float32."""
    print(f'frozen swe {gen} {dtype} reorder={reorder}', flush=True)
    mesh = wmesh.tri2(64, jitter=0.2, seed=3) if gen == 'tri2_jitter' else \
        getattr(wmesh, gen)(64)
    geo = oswe.geometry(mesh, centre=(0.45, 0.55), width=200.0, amp=0.1)
    dt = wswe.cfl_dt(mesh, geo['h'], 9.81)
    eqn = wswe.ShallowWaterEquation(mesh, geo, dt=dt, g=9.81, dtype=dtype)
    esr.compile([eqn], 'hip', freeze=True, reorder=reorder)
    luts = sum(len(k.luts) + len(k.consts) for k in eqn.easier_engine.kernels)
    if reorder == 'blocks':               # the gathers go through LDS halos
        assert 'H1_0' in eqn.easier_engine.source
    assert (luts > 0) == (gen != 'tri2_jitter') or dtype == torch.float32
    ref = oswe.SWE(mesh, geo, dt=dt, g=9.81, dtype=dtype)
    for _ in range(30):
        eqn()
        ref()
    for k in ('h', 'uh', 'vh'):
        got, want = getattr(eqn, k).collect(), getattr(ref, k)
        if dtype == torch.float64:
            assert torch.equal(got, want), k
        else:
            assert float((got - want).abs().max()) <= 1e-5 * float(want.abs().max()), k


@pytest.mark.parametrize('option', [('rows_per_thread', 2), ('pack_intermediates', True),
                                    ('waves_per_eu', 8), ('remat', False),
                                    ('late_eval', False)])
@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_codegen_options_are_bit_identical(gpu, dtype, option):
    """Fused kernels generated with two rows per thread (vector row
    accesses), with row-interleaved intermediate buffers, with an occupancy
    request, with gathered stage inputs written instead of recomputed, or with
    every value in its earliest kernel give the same bits as the defaults: the SWE step (even cell count) and a program on an odd
    number of rows (the last thread's second row is past the end)."""
    from easier_amd.core import fuser
    key, val = option
    print(f'codegen {key}={val} {dtype}', flush=True)
    mesh = wmesh.tri2(10)
    geo = oswe.geometry(mesh, centre=(0.45, 0.55), width=200.0, amp=0.1)
    dt = 0.2 * wswe.cfl_dt(mesh, geo['h'], 9.81)

    class Odd(esr.Module):
        def __init__(self, n=1001):
            super().__init__()
            g = torch.Generator().manual_seed(9)
            src = torch.randint(0, n, (3 * n,), generator=g)
            dst = torch.arange(3 * n) % n
            self.red = esr.Reducer(dst, n)
            self.sel = esr.Selector(src)
            self.u = esr.Tensor(torch.rand(n, generator=g, dtype=torch.float64).to(dtype),
                                mode='partition')
            self.w = esr.Tensor(torch.rand(3 * n, generator=g, dtype=torch.float64).to(dtype),
                                mode='partition')

        def forward(self):
            self.u[:] = self.u * 0.5 + self.red(self.sel(self.u) * self.w)

    outs = []
    saved = dict(fuser.OPTIONS)
    try:
        for v in (saved[key], val):
            fuser.OPTIONS[key] = v
            eqn = wswe.ShallowWaterEquation(mesh, geo, dt=dt, g=9.81, dtype=dtype)
            odd = Odd()
            esr.compile([eqn, odd], 'hip')
            if key == 'rows_per_thread':
                assert all(k.rpt == v for k in eqn.easier_engine.kernels)
            for _ in range(5):
                eqn()
                odd()
            outs.append([eqn.h.collect(), eqn.uh.collect(), eqn.vh.collect(), odd.u.collect()])
    finally:
        fuser.OPTIONS.update(saved)
    for a, b in zip(*outs):
        assert torch.equal(a, b)
    if dtype == torch.float64:
        ref = oswe.SWE(mesh, geo, dt=dt, g=9.81, dtype=dtype)
        for _ in range(5):
            ref()
        assert torch.equal(outs[1][0], ref.h)


@pytest.mark.parametrize('reorder,freeze', [('rcm', False), ('blocks', True)])
def test_compile_cache_round_trip(gpu, tmp_path, reorder, freeze):
    """esr.dump() then esr.compile(load_dir=...) on a fresh program: the same
    renumbering, and the SWE steps of both compiles equal the oracle's bit
    for bit (float64)."""
    mesh = wmesh.tri2(12) if reorder == 'rcm' else wmesh.tri4(24)
    geo = oswe.geometry(mesh, centre=(0.45, 0.55), width=200.0, amp=0.1)
    dt = 0.2 * wswe.cfl_dt(mesh, geo['h'], 9.81)
    ref = oswe.SWE(mesh, geo, dt=dt, g=9.81)
    for _ in range(5):
        ref()
    states = []
    for load in (None, str(tmp_path)):
        eqn = wswe.ShallowWaterEquation(mesh, geo, dt=dt, g=9.81)
        esr.compile([eqn], 'hip', load_dir=load, reorder=reorder, freeze=freeze)
        if load is None:
            esr.dump([eqn], str(tmp_path))
        for _ in range(5):
            eqn()
        states.append([eqn.h.collect(), eqn.uh.collect(), eqn.vh.collect(),
                       {eqn.easier_program.space_name[k]: v
                        for k, v in eqn.easier_program.space_order.items()}])
    for a, b, k in zip(states[0][:3], states[1][:3], ('h', 'uh', 'vh')):
        assert torch.equal(a, b)
        assert torch.equal(a, getattr(ref, k)), k
    assert states[0][3].keys() == states[1][3].keys()
    for k in states[0][3]:
        assert np.array_equal(states[0][3][k], states[1][3][k])


def test_fused_cg_graph_replays_are_deterministic(gpu):
    eqn = Poisson5(40, seed=3)
    sol = CG(eqn.A, eqn.b, eqn.x, Jacobi(eqn.Ac))
    esr.compile([eqn, sol], 'hip')
    assert sol.easier_fused is not None
    info1 = sol.solve(rtol=1e-9, maxiter=5000)
    x1 = eqn.x.collect()
    ops_zero = torch.zeros_like(eqn.x.data)
    eqn.x.data.copy_(ops_zero)
    info2 = sol.solve(rtol=1e-9, maxiter=5000)
    assert info1 == info2
    assert torch.equal(x1, eqn.x.collect())
    # the generic compiled modules agree with the fused iteration
    eqn.x.data.copy_(ops_zero)
    fused, sol.easier_fused = sol.easier_fused, None
    info3 = sol.solve(rtol=1e-9, maxiter=5000)
    sol.easier_fused = fused
    assert abs(info3['iters'] - info1['iters']) <= 1
    torch.testing.assert_close(eqn.x.collect(), x1, rtol=1e-8, atol=1e-10)


def test_mixed_precision_cg_matches_fp64_oracle(gpu):
    from easier_amd.numeric.solver import MixedPrecisionCG
    from easier_amd.workloads import poisson as wp
    mesh, sys_, _ = wp.varcoef_system(24, seed=5)
    eqn = wp.PoissonP1(sys_)
    sol = MixedPrecisionCG(eqn.A, eqn.b, eqn.x, eqn.M)
    esr.compile([eqn, sol], 'hip')
    info = sol.solve(rtol=1e-10, inner_rtol=1e-4, maxiter=10000)
    assert info['outer'] >= 2
    assert info['residual'] <= 1e-10 * info['bnorm']
    A = osol.Linsys(torch.from_numpy(sys_['Av']), torch.from_numpy(sys_['Ae']),
                    torch.from_numpy(sys_['col']), torch.from_numpy(sys_['row']),
                    sys_['Av'].shape[0])
    b = torch.from_numpy(sys_['b'])
    x_o = torch.zeros_like(b)
    osol.cg(A, b, x_o, M=lambda r: r / A.Av, rtol=1e-12, maxiter=20000)
    x = eqn.x.collect()
    err = float((x - x_o).norm() / x_o.norm())
    assert err < 1e-7, err
    # the true residual the solver reports is the float64 one
    r = b - A(x)
    torch.testing.assert_close(torch.tensor(info['residual'], dtype=torch.float64), r.norm(),
                               rtol=1e-6, atol=0)


def test_recomputed_gathers_match_the_node_engine(gpu):
    """Gathered elementwise values recomputed at their source rows (random
    Selector indices, so neighbours lie anywhere), a gathered value over a
    tensor its gathering kernel reassigns (written and read back instead),
    and a partial sum evaluated in a later kernel: the fused kernels give the
    node-by-node engine's bits over several forwards."""
    from easier_amd.core import fuser
    n = 3001

    class Prog(esr.Module):
        def __init__(self):
            super().__init__()
            g = torch.Generator().manual_seed(11)
            self.sel = esr.Selector(torch.randint(0, n, (3 * n,), generator=g))
            self.red = esr.Reducer(torch.arange(3 * n) % n, n)
            mk = lambda: esr.Tensor(torch.rand(n, dtype=torch.float64, generator=g),
                                    mode='partition')
            self.u, self.v, self.w = mk(), mk(), mk()
            self.c = esr.Tensor(torch.rand(3 * n, dtype=torch.float64, generator=g),
                                mode='partition')
            self.x = esr.Tensor(torch.zeros(n, dtype=torch.float64), mode='partition')

        def forward(self):
            d1 = self.red(self.sel(self.w * 0.5 + self.v) * self.c)
            d2 = self.red(self.sel(self.w + 0.25 * d1))
            d3 = self.red(self.sel(self.w - 0.125 * d2))
            self.x[:] = self.x + (d1 + d2 + d3) * 0.01
            self.u[:] = self.red(self.sel(self.u * 3)) * 0.25

    outs = []
    for fuse in (True, False):
        m = Prog()
        print(f'recomputed gathers, fuse={fuse}', flush=True)
        esr.compile([m], 'hip', fuse=fuse)
        eng = m.easier_engine
        if fuse:
            assert isinstance(eng, fuser.FusedEngine)
        for _ in range(4):
            m()
        outs.append((m.x.collect(), m.u.collect()))
    assert torch.equal(outs[0][0], outs[1][0])
    assert torch.equal(outs[0][1], outs[1][1])


def test_replaced_storage_is_refused(gpu):
    """Compiled kernels and captured graphs hold device addresses: giving a
    compiled tensor new storage (t.data = ...) after compile() raises at the
    next launch instead of computing on the old storage; in-place updates
    (t.data.copy_) are seen."""
    eqn = Poisson5(12, seed=1)
    sol = CG(eqn.A, eqn.b, eqn.x, Jacobi(eqn.Ac))
    esr.compile([eqn, sol], 'hip')
    assert sol.easier_fused is not None
    sol.solve(rtol=1e-8, maxiter=500)
    eqn.x.data.copy_(torch.zeros_like(eqn.x.data))
    info = sol.solve(rtol=1e-8, maxiter=500)
    assert info['residual'] <= 1e-8 * float(eqn.b.data.norm())
    old = eqn.x.data
    eqn.x.data = torch.zeros_like(old)
    with pytest.raises(RuntimeError, match='storage of a compiled tensor'):
        sol.solve(rtol=1e-8, maxiter=500)
    eqn.x.data = old
    sol.solve(rtol=1e-8, maxiter=500)

    mesh = wmesh.tri2(8)
    geo = oswe.geometry(mesh, centre=(0.45, 0.55), width=200.0, amp=0.1)
    swe = wswe.ShallowWaterEquation(mesh, geo, dt=0.2 * wswe.cfl_dt(mesh, geo['h'], 9.81),
                                    g=9.81)
    esr.compile([swe], 'hip')
    assert type(swe.easier_engine).__name__ == 'FusedEngine'
    swe()
    swe()                                   # now a graph replay
    h = swe.h.data
    swe.h.data = h.clone()
    with pytest.raises(RuntimeError, match='storage of a compiled tensor'):
        swe()
    swe.h.data = h
    swe()


def test_hdf5_loaders_and_tensor_save(gpu, tmp_path):
    """A program whose Selector, Reducer and Tensor come from esr.hdf5
    datasets of a mesh file (the reference's tutorial layout) computes what
    the same program built from in-memory arrays computes; Tensor.save adds
    datasets that read back as collect()."""
    from easier_amd.core import h5
    mesh = wmesh.tri4(8)
    path = str(tmp_path / 'mesh.h5')
    wmesh.save_h5(mesh, path)
    nc = mesh['cells'].shape[0]
    g = torch.Generator().manual_seed(11)
    w = torch.randn(mesh['src'].shape[0], generator=g, dtype=torch.double)
    h5.append(path, {'w': w.numpy()})

    class P(esr.Module):
        def __init__(self, from_file):
            super().__init__()
            if from_file:
                src, dst = esr.hdf5(path, 'src'), esr.hdf5(path, 'dst')
                wt = esr.hdf5(path, 'w')
            else:
                src, dst, wt = (torch.from_numpy(mesh['src']), torch.from_numpy(mesh['dst']), w)
            self.sel = esr.Selector(src)
            self.red = esr.Reducer(dst, nc)
            self.w = esr.Tensor(wt, mode='partition')
            self.v = esr.Tensor(torch.arange(nc, dtype=torch.double) / nc, mode='partition')
            self.out = esr.Tensor(esr.zeros((nc,), dtype=torch.double), mode='partition')

        def forward(self):
            self.out[:] = self.red(self.sel(self.v) * self.w) + self.v

    a, b = P(True), P(False)
    esr.compile([a], 'hip')
    esr.compile([b], 'hip')
    a()
    b()
    ref = oops.reducer(oops.selector(torch.arange(nc, dtype=torch.double) / nc,
                                     torch.from_numpy(mesh['src'])) * w,
                       torch.from_numpy(mesh['dst']), nc) + torch.arange(nc, dtype=torch.double) / nc
    assert torch.equal(a.out.collect(), b.out.collect())
    assert torch.equal(a.out.collect(), ref)
    out_path = str(tmp_path / 'run' / 'out.h5')
    a.out.save(out_path, 'fields/out')
    a.v.save(out_path, 'fields/v')
    assert np.array_equal(h5.read(out_path, 'fields/out'), a.out.collect().numpy())
    assert np.array_equal(h5.read(out_path, 'fields/v'), a.v.collect().numpy())
    with pytest.raises(ValueError, match='exists'):
        a.out.save(out_path, 'fields/out')
