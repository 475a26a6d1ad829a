"""Benchmark: shallow-water RK4 steps on one MI355X (the hot path of
esr.Module forwards: cell->edge Selector, flux, edge->cell Reducer, update).

Workload (default): the 2-triangle mesh of a 2048 x 2048 cell grid
(8,388,608 triangles, 25,157,632 directed cell-to-cell edges, 8,192 wall
faces), state (h, hu, hv) in float64, h = 1 + 0.1 exp(-|x - x0|^2 / 0.005)
with x0 ~ U[0.3, 0.7]^2 (seed 0), hu = hv = 0, g = 9.81, reflective walls,
dt from CFL 0.4. The program is the tutorial's ShallowWaterEquation module
compiled with esr.compile(..., 'hip', freeze=True) (the geometry, which no
forward writes, may be read through bit-exact dictionaries; --no-freeze
reads it as plain arrays, and the line carries that step time too); a step
is one forward() (one RK4 step).

Prints one JSON line (rank 0, N = 1). Side figures: the P1 Poisson
Jacobi-PCG solve of the 100 x 100 mesh (10,201 vertices, tol 1e-8), the
primitives, the float32 step and, by default (--configs 3,4,5), the batched
Jacobi-PCG solves (64 right-hand sides on the 1024 x 1024 mesh, 256 meshes
of sides 32..512), the variable-coefficient 2048 x 2048 Poisson solves:
float64 and mixed-precision Jacobi-PCG and multigrid flexible CG to 1e-10,
and Jacobi-PCG against the multigrid and learned-multigrid flexible CG to
1e-8 (the metric's solve-time half).
"""
import argparse
import json
import math
import os
import platform
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = ("Poisson solve time to 1e-8 rel. residual (learned vs Jacobi PCG); "
          "SWE cells/s")
HBM_PEAK = 8.0e12


def log(msg):
    print(f'[bench {time.strftime("%H:%M:%S")}] {msg}', file=sys.stderr, flush=True)


def swe_bytes(nc, ne, nbc, T, I=4):
    """Algorithmic bytes of one RK4 step by the per-unit figure of the
    reference's data layout: per stage kernel, per directed edge the source
    index and (sx, sy, alpha), per wall face its index and (bsx, bsy); the 55
    cell-field reads/writes of a step split over the 4 stages as 10 + 13 + 13
    + 19 (state, area, stage inputs, stage deltas, update):
    4 [ne (I + 3T) + nbc (I + 2T)] + 55 nc T per step (6.51 GB in float64 on
    the 2048 x 2048 mesh)."""
    per_stage_edges = ne * (I + 3 * T) + nbc * (I + 2 * T)
    cells = {0: 10, 1: 13, 2: 13, 3: 19}
    stages = [per_stage_edges + cells[k] * nc * T for k in range(4)]
    return stages, sum(stages)


def build_swe(n, dtype, seed, fuse=True, freeze=False, reorder='rcm'):
    import numpy as np
    import torch
    import easier_amd as esr
    from easier_amd.workloads import mesh as wmesh
    from easier_amd.workloads import swe as wswe
    t0 = time.perf_counter()
    mesh = wmesh.tri2(n)
    log(f'mesh {n}x{n}: {mesh["cells"].shape[0]} cells, {mesh["src"].shape[0]} '
        f'edges ({time.perf_counter() - t0:.1f} s)')
    x0 = wswe.gaussian_centre(seed)
    t0 = time.perf_counter()
    geo = wswe.assemble_geometry(mesh, centre=x0, width=1.0 / 0.005, amp=0.1)
    log(f'geometry programs on the GPU ({time.perf_counter() - t0:.1f} s)')
    g = 9.81
    dt = wswe.cfl_dt(mesh, geo['h'], g, cfl=0.4)
    t0 = time.perf_counter()
    eqn = wswe.ShallowWaterEquation(mesh, geo, dt=dt, g=g, dtype=dtype)
    esr.compile([eqn], 'hip', fuse=fuse, freeze=freeze, reorder=reorder)
    torch.cuda.synchronize()
    log(f'compiled ({type(eqn.easier_engine).__name__}, '
        f'{time.perf_counter() - t0:.1f} s), dt = {dt:.3e}')
    return mesh, geo, eqn, dict(x0=x0, g=g, dt=dt)


def time_kernels(eqn, reps):
    """Average duration of each fused kernel, from HIP events on the stream
    the kernels are launched on (the current stream): `reps` back-to-back
    launches of the one kernel between two events, so no event sits between
    launches. (A stage kernel relaunched reads the arrays it reads inside
    the step; the last one updates the state in place, which leaves the
    state advanced but the work per launch unchanged.)"""
    import ctypes
    import torch
    from easier_amd import _native as nat
    eng = eqn.easier_engine
    if not hasattr(eng, 'launches'):
        return None
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    out = []
    for fn, grid, buf, block in eng.launches:
        nat.call('esr_launch', fn, grid, 1, 1, block, 1, 1, 0, sp,
                 ctypes.addressof(buf), ctypes.sizeof(buf))
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            nat.call('esr_launch', fn, grid, 1, 1, block, 1, 1, 0, sp,
                     ctypes.addressof(buf), ctypes.sizeof(buf))
        b.record(stream)
        torch.cuda.synchronize()
        out.append(a.elapsed_time(b) / 1e3 / reps)
    return out


def host_cpu():
    """(model name from /proc/cpuinfo, CPUs this process may run on)."""
    model = platform.processor() or platform.machine()
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return model, cores


def cpu_baseline_swe(mesh, geo, meta, dtype, steps, samples=3):
    """The oracle (torch CPU operators of the reference's CPU backend) on the
    same mesh and state: after one warm-up step, the median of `samples`
    timed runs of `steps` RK4 steps, on every CPU this process may use."""
    import torch
    from oracle import swe as oswe
    model, cores = host_cpu()
    torch.set_num_threads(cores)
    ref = oswe.SWE(mesh, geo, dt=meta['dt'], g=meta['g'], dtype=dtype)
    ref()                                   # warm-up
    times = []
    for _ in range(samples):
        t0 = time.perf_counter()
        for _ in range(steps):
            ref()
        times.append(time.perf_counter() - t0)
    dt = sorted(times)[len(times) // 2]
    nc = mesh['cells'].shape[0]
    return dict(value=nc * steps / dt, unit='cells/s', cores=torch.get_num_threads(),
                kind='port', sample=f'median of {samples} runs of {steps} RK4 steps of '
                f'the same {nc}-cell mesh after 1 warm-up step, torch CPU operators',
                seconds=dt, all_seconds=times, cpu_model=model,
                os_cpu_count=os.cpu_count())


def pcg_iteration_roofline(cg, n, nnz, T, chunks=3):
    """One Jacobi-PCG iteration of a fused CG, timed with HIP events on the
    launch stream around `chunks` replays of its captured chunk graph
    (32 iterations each) of a fresh solve armed with tolerance 0, so that
    every iteration runs. Algorithmic bytes of an iteration:
    nnz (I + T) + (n + 1) I + 14 n T (the operator once, 14 vector passes)."""
    import torch
    f = cg.easier_fused
    cg.x.data.zero_()
    f.arm(0.0, None, None)
    stream = torch.cuda.current_stream()
    f.graph.replay()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(chunks):
        f.graph.replay()
    b.record(stream)
    torch.cuda.synchronize()
    it = a.elapsed_time(b) / 1e3 / (chunks * f.graph_chunk)
    bytes_ = nnz * (4 + T) + (n + 1) * 4 + 14 * n * T
    f = cg.easier_fused
    k1 = 'k_pcg_step1_sell' if getattr(f, 'sell', None) is not None else 'k_pcg_step1'
    return dict(bound='hbm', kernel=f'Jacobi-PCG iteration ({k1}, k_pcg_step2, k_pcg_update '
                                    'of one step)',
                achieved=bytes_ / it / 1e9, peak=HBM_PEAK / 1e9, unit='GB/s',
                frac=bytes_ / it / HBM_PEAK, traffic=None, bytes_per_launch=bytes_,
                avg_iteration_s=it)


def committed_pcg_traffic(dtype):
    """HBM bytes of one Jacobi-PCG iteration (its three kernels) on the
    2048 x 2048 variable-coefficient system from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary in profiles/ (line 'traffic_bytes iteration <bytes>'),
    or None."""
    path = os.path.join(REPO, 'profiles', f'pcg_varcoef_2048_{dtype}_traffic.md')
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[:2] == ['traffic_bytes', 'iteration']:
                return int(parts[2])
    return None


def pcg_profile(dtype):
    """The Jacobi-PCG iteration alone: the 2048 x 2048 variable-coefficient
    system (seed 0, the mixed-precision workload's) compiled in `dtype`,
    pcg_iteration_roofline over 3 chunks."""
    import torch
    import easier_amd as esr
    from easier_amd.numeric.solver import CG
    from easier_amd.workloads import poisson as wp
    _, sys_, _ = wp.varcoef_system(2048, seed=0)
    e1 = wp.PoissonP1(sys_, dtype=torch.float64 if dtype == 'f64' else torch.float32)
    cg = CG(e1.A, e1.b, e1.x, e1.M)
    esr.compile([e1, cg], 'hip')
    T = 8 if dtype == 'f64' else 4
    roof = pcg_iteration_roofline(cg, int(sys_['Av'].shape[0]), int(sys_['row'].shape[0]), T)
    roof['traffic'] = committed_pcg_traffic(dtype)
    torch.cuda.synchronize()
    return dict(workload=f'Jacobi-PCG iteration, variable-coefficient P1 Poisson 2048x2048 '
                f'(4,198,401 vertices), {dtype}', roofline=roof)


def poisson_config1(skip_cpu):
    import torch
    import easier_amd as esr
    from easier_amd.numeric.solver import CG
    from easier_amd.workloads import poisson as wp
    mesh, sys_ = wp.config1_system(100, seed=0)
    eqn = wp.PoissonP1(sys_)
    sol = CG(eqn.A, eqn.b, eqn.x, eqn.M)
    esr.compile([eqn, sol], 'hip')
    res = None
    times = []
    for rep in range(4):
        torch.cuda.synchronize()
        eqn.x.data.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = sol.solve(rtol=1e-8, maxiter=5000)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    out = dict(workload='P1 Poisson 100x100 (10,201 vertices), Jacobi-PCG, FP64, '
               'rtol 1e-8', solve_ms=1e3 * sorted(times[1:])[1], iters=res['iters'],
               residual=res['residual'])
    if not skip_cpu:
        from oracle import solvers as osol
        A = osol.Linsys(torch.from_numpy(sys_['Av']), torch.from_numpy(sys_['Ae']),
                        torch.from_numpy(sys_['col']), torch.from_numpy(sys_['row']),
                        sys_['Av'].shape[0])
        b = torch.from_numpy(sys_['b'])
        ct = []
        for rep in range(3):
            x = torch.zeros_like(b)
            t0 = time.perf_counter()
            info, _ = osol.cg(A, b, x, M=lambda r: r / A.Av, rtol=1e-8, maxiter=5000)
            ct.append(time.perf_counter() - t0)
        out['cpu_solve_ms'] = 1e3 * sorted(ct)[1]
        out['cpu_iters'] = info['iters']
    return out


def primitives(eqn, reps=20, with_cpu=False):
    """The plain Selector / Reducer kernels on the compiled SWE mesh: the
    cell -> edge gather of `gather_src` (esr_gather) and the edge -> cell CSR
    sum of `scatter` (esr_csr_reduce), each timed with HIP events over `reps`
    back-to-back launches on the launch stream; algorithmic bytes
    n_out (I + T) + n_src T and n_in T + (n_out + 1) I + n_out T."""
    import torch
    from easier_amd.core import ops
    stream = torch.cuda.current_stream()
    src = eqn.h.data
    sel, red = eqn.gather_src, eqn.scatter
    T, I = src.element_size(), 4
    edge_vals = ops.gather(src, sel.idx)
    out_cells = torch.empty_like(src)

    def run(fn):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / 1e3 / reps

    t_g = run(lambda: ops.gather(src, sel.idx, out=edge_vals))
    t_r = run(lambda: ops.csr_reduce(edge_vals, red.easier_rowptr, red.easier_perm,
                                     red.easier_long_rows, red.n, 'sum', out=out_cells))
    # aggregators (esr.sum / esr.norm of a cell tensor): two fixed-shape stages
    col = src.reshape(-1, 1)
    agg_out = torch.empty(1, dtype=src.dtype, device=src.device)
    ws = torch.empty(ops.aggregate_ws_bytes(col), dtype=torch.uint8, device=src.device)
    t_s = run(lambda: ops.aggregate('sum', col, out=agg_out, ws=ws))
    t_n = run(lambda: ops.aggregate('norm', col, out=agg_out, ws=ws))
    ne, nc = sel.idx.shape[0], src.shape[0]
    bg = ne * (I + T) + nc * T
    br = ne * T + (nc + 1) * I + nc * T
    ba = nc * T
    cpu = {}
    if with_cpu:
        # the oracle's operators on the host (torch CPU, all threads): t[idx]
        # and zeros().scatter_reduce_(..., include_self=False), best of 3
        from oracle import ops as oops
        hsrc = src.cpu()
        hidx = sel.idx.cpu().long()
        hvals = edge_vals.cpu()
        ncells = src.shape[0]
        hrow = torch.repeat_interleave(torch.arange(ncells),
                                       torch.diff(red.easier_rowptr.cpu().long()))

        def host(fn):
            best = None
            for _ in range(3):
                t0 = time.perf_counter()
                fn()
                dt = time.perf_counter() - t0
                best = dt if best is None else min(best, dt)
            return best
        cpu['gather'] = host(lambda: oops.selector(hsrc, hidx))
        cpu['reduce'] = host(lambda: oops.reducer(hvals, hrow, ncells))
    return {
        'selector_gather': dict(edges=ne, seconds=t_g, bytes=bg, GBps=bg / t_g / 1e9,
                                frac=bg / t_g / HBM_PEAK,
                                cpu_oracle_seconds=cpu.get('gather')),
        'reducer_csr_sum': dict(rows=nc, entries=ne, seconds=t_r, bytes=br,
                                GBps=br / t_r / 1e9, frac=br / t_r / HBM_PEAK,
                                cpu_oracle_seconds=cpu.get('reduce')),
        'aggregate_sum': dict(n=nc, seconds=t_s, bytes=ba, GBps=ba / t_s / 1e9,
                              frac=ba / t_s / HBM_PEAK),
        'aggregate_norm': dict(n=nc, seconds=t_n, bytes=ba, GBps=ba / t_n / 1e9,
                               frac=ba / t_n / HBM_PEAK),
    }


def poisson_gmres_config1(skip_cpu):
    """GMRES(20), Jacobi-preconditioned, on the config-1 Poisson system:
    the fused Arnoldi cycles against the oracle's GMRES on the host."""
    import torch
    import easier_amd as esr
    from easier_amd.numeric.solver import GMRES
    from easier_amd.workloads import poisson as wp
    mesh, sys_ = wp.config1_system(100, seed=0)
    eqn = wp.PoissonP1(sys_)
    sol = GMRES(eqn.A, eqn.b, eqn.x, eqn.M, restart=20)
    esr.compile([eqn, sol], 'hip')
    times, res = [], None
    for rep in range(4):
        eqn.x.data.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = sol.solve(rtol=1e-8, maxiter=500)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    out = dict(workload='P1 Poisson 100x100 (10,201 vertices), Jacobi-preconditioned '
               'GMRES(20), FP64, rtol 1e-8', solve_ms=1e3 * sorted(times[1:])[1],
               cycles=res['iters'], residual=res['residual'])
    if not skip_cpu:
        from oracle import solvers as osol
        A = osol.Linsys(torch.from_numpy(sys_['Av']), torch.from_numpy(sys_['Ae']),
                        torch.from_numpy(sys_['col']), torch.from_numpy(sys_['row']),
                        sys_['Av'].shape[0])
        b = torch.from_numpy(sys_['b'])
        x = torch.zeros_like(b)
        t0 = time.perf_counter()
        info, _ = osol.gmres(A, b, x, M=lambda r: r / A.Av, restart=20, rtol=1e-8,
                             maxiter=500)
        out['cpu_solve_ms'] = 1e3 * (time.perf_counter() - t0)
        out['cpu_cycles'] = info['iters']
    return out


def swe_side(n, seed, dtype, freeze, steps=100, warmup=5, with_primitives=False):
    """The SWE step in another precision or compile mode (cells/s); with
    with_primitives, also the plain Selector / Reducer / aggregator kernels
    on this compile's mesh order (RCM when not frozen)."""
    import torch
    mesh, geo, eqn, meta = build_swe(n, dtype, seed, freeze=freeze,
                                     reorder='blocks' if freeze else 'rcm')
    for _ in range(warmup):
        eqn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eqn()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    nc = mesh['cells'].shape[0]
    out = dict(workload=f'shallow-water RK4 {n}x{n}, {dtype}, freeze={freeze}', steps=steps,
               ms_per_step=1e3 * wall / steps, cells_per_s=nc * steps / wall,
               state_finite=bool(torch.isfinite(eqn.h.collect()).all()))
    if with_primitives:
        out['primitives'] = primitives(eqn)
        out['primitives']['mesh_order'] = 'rcm' if not freeze else 'blocks'
    return out


def committed_traffic(n, dtype, kernel, tag=''):
    """HBM bytes per launch of `kernel` measured by the rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes summarised in profiles/ (lines 'traffic_bytes <kernel>
    <bytes>'), or None when no summary covers this workload."""
    path = os.path.join(REPO, 'profiles', f'swe_fused_{n}_{dtype}{tag}_traffic.md')
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] == 'traffic_bytes' and parts[1] == kernel:
                return int(parts[2])
    return None


def _best_time(fn, reps, reset=None):
    import torch
    best, out = None, None
    for rep in range(reps + 1):                 # first call warms up
        if reset is not None:
            reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        if rep > 0:
            best = dt if best is None else min(best, dt)
    return best, out


def batched_config3(reps=1):
    """(a) 64 right-hand sides U[-1,1] (seeds 0..63) on the 1024 x 1024 mesh;
    (b) 256 meshes, n ~ U{32..512}, interior vertices jittered by U[-0.2h,
    0.2h], f ~ U[-1,1] each: Jacobi-PCG FP64 to 1e-8 per instance, one batched
    solve (autotuned tile height) against the instances solved one after the
    other with the single-system fused solver."""
    import numpy as np
    from easier_amd.numeric.solver.batched import BatchedCG
    from easier_amd.workloads import mesh as wmesh
    from easier_amd.workloads import poisson as wp
    out = {}
    t0 = time.perf_counter()
    m = wmesh.tri2(1024)
    base = wp.p1_system(m, None, None)
    nv = m['points'].shape[0]
    # the 64 load vectors of the one mesh (the assembler's bits, one pass)
    loads = wp.load_vectors(m, [wp.rhs_uniform(nv, k) for k in range(64)], base['boundary'])
    sys_a = [dict(base, b=loads[k]) for k in range(64)]
    rng = np.random.default_rng(0)
    sys_b = []
    for k in range(256):
        n = int(rng.integers(32, 513))
        mk = wmesh.tri2(n, jitter=0.2, seed=k)
        sys_b.append(wp.p1_system(mk, None, rng.uniform(-1, 1, mk['points'].shape[0])))
    log(f'config 3 systems built ({time.perf_counter() - t0:.1f} s)')
    for name, systems in (('3a_64rhs_1024', sys_a), ('3b_256meshes', sys_b)):
        bcg = BatchedCG(systems)
        R, timings = bcg.autotune(iters=5)
        t, infos = _best_time(lambda: bcg.solve(rtol=1e-8, maxiter=20000), reps)
        its = [i['iters'] for i in infos]
        ok = all(i['residual'] <= 1e-8 * float(np.linalg.norm(s['b'])) * (1 + 1e-12)
                 for i, s in zip(infos, systems))
        ts = 0.0
        for s in systems[:8]:                # one at a time, on a sample
            single = BatchedCG([s])
            single.R = R or 4
            dt, _ = _best_time(lambda: single.solve(rtol=1e-8, maxiter=20000), 1)
            ts += dt
        out[name] = dict(instances=len(systems), rows=int(bcg.offsets[-1]),
                         solver='shared operator' if bcg.multi is not None else 'batched',
                         tile_rows=None if R is None else 256 * R, autotune_ms=timings,
                         solve_s=t, iters_max=max(its), iters_mean=float(np.mean(its)),
                         all_converged=ok, sequential_first8_s=ts,
                         sequential_all_est_s=ts * len(systems) / 8)
        log(f'config {name}: {t:.3f} s batched, iters max {max(its)}')
        del bcg
    return out


def cpu_pcg_rate(sys_, iters=20):
    """Seconds per Jacobi-PCG iteration of the oracle (torch CPU operators of
    the reference's CPU backend, all host threads) on a system, from a
    bounded run of `iters` iterations after one warm-up iteration."""
    import torch
    from oracle import solvers as osol
    A = osol.Linsys(torch.from_numpy(sys_['Av']), torch.from_numpy(sys_['Ae']),
                    torch.from_numpy(sys_['col']), torch.from_numpy(sys_['row']),
                    sys_['Av'].shape[0])
    b = torch.from_numpy(sys_['b'])
    osol.cg(A, b, torch.zeros_like(b), M=lambda r: r / A.Av, rtol=0.0, maxiter=1)
    t0 = time.perf_counter()
    osol.cg(A, b, torch.zeros_like(b), M=lambda r: r / A.Av, rtol=0.0, maxiter=iters)
    return (time.perf_counter() - t0) / (iters + 1)


def mixed_config4(reps=1, skip_cpu=False):
    """Variable-coefficient Poisson on the 2048 x 2048 mesh (4,198,401
    vertices): float64 Jacobi-PCG to 1e-10 against the mixed-precision solver
    (float32 inner CG, float64 refinement) to 1e-10; error of the mixed
    solution against the float64 one."""
    import numpy as np
    import easier_amd as esr
    from easier_amd.numeric.solver import CG, MixedPrecisionCG
    from easier_amd.workloads import poisson as wp
    t0 = time.perf_counter()
    mesh, sys_, _ = wp.varcoef_system(2048, seed=0)
    log(f'config 4 system built ({time.perf_counter() - t0:.1f} s)')
    e1 = wp.PoissonP1(sys_)
    cg = CG(e1.A, e1.b, e1.x, e1.M)
    esr.compile([e1, cg], 'hip')
    t64, i64 = _best_time(lambda: cg.solve(rtol=1e-10, maxiter=100000), reps,
                          reset=lambda: e1.x.data.zero_())
    x64 = e1.x.collect().numpy()
    roof = pcg_iteration_roofline(cg, int(sys_['Av'].shape[0]), int(sys_['row'].shape[0]), 8)
    roof['traffic'] = committed_pcg_traffic('f64')
    log(f'config 4 fp64 Jacobi-PCG iteration: {roof["avg_iteration_s"] * 1e6:.1f} us, '
        f'{roof["frac"]:.1%} of 8 TB/s')
    e2 = wp.PoissonP1(sys_)
    mp = MixedPrecisionCG(e2.A, e2.b, e2.x, e2.M)
    esr.compile([e2, mp], 'hip')
    tmx, imx = _best_time(lambda: mp.solve(rtol=1e-10, inner_rtol=1e-4, maxiter=100000),
                          reps, reset=lambda: e2.x.data.zero_())
    xmx = e2.x.collect().numpy()
    out = dict(roofline=roof, fp64_s_per_iter=t64 / max(i64['iters'] + 1, 1),
               mixed_s_per_inner_iter=tmx / max(imx['iters'], 1),
               vertices=int(sys_['Av'].shape[0]), fp64_s=t64, fp64_iters=i64['iters'],
               fp64_residual=i64['residual'], mixed_s=tmx, mixed_inner_iters=imx['iters'],
               mixed_outer=imx['outer'], mixed_true_residual=imx['residual'],
               bnorm=imx['bnorm'],
               rel_err_vs_fp64=float(np.linalg.norm(xmx - x64) / np.linalg.norm(x64)))
    log(f'config 4: fp64 {t64:.3f} s ({i64["iters"]} it), mixed {tmx:.3f} s')
    if not skip_cpu:
        rate = cpu_pcg_rate(sys_)
        out['cpu_fp64_s_per_iter'] = rate
        out['cpu_fp64_solve_estimate_s'] = rate * i64['iters']
        out['cpu_sample'] = ('21 Jacobi-PCG iterations of the oracle (torch CPU) on the '
                             'same system; estimate = s/iter x the float64 iteration count')
        log(f'config 4 CPU: {rate:.3f} s/iter')
    # the same problem with the multigrid preconditioner in flexible CG:
    # float64 cycle, and float32 cycle (float64 dots, iteration, residual)
    import torch
    from easier_amd.numeric.preconditioner.multigrid import MultigridPreconditioner
    from easier_amd.numeric.solver import FlexibleCG
    logk = wp.gaussian_random_field(2048, 0).reshape(-1)
    for tag, cd in (('mg_fp64', torch.float64), ('mg_fp32_cycle', torch.float32)):
        e3 = wp.PoissonP1(sys_)
        fc = FlexibleCG(e3.A, e3.b, e3.x, MultigridPreconditioner(None, cycle_dtype=cd),
                        dict(points=mesh['points'], logk=logk, n_grid=2048))
        esr.compile([e3, fc], 'hip')
        tm, im = _best_time(lambda: fc.solve(rtol=1e-10, maxiter=1000), max(reps, 2),
                            reset=lambda: e3.x.data.zero_())
        xm = e3.x.collect().numpy()
        out[tag] = dict(seconds=tm, iters=im['iters'], true_residual=im['residual'],
                        rel_err_vs_fp64=float(np.linalg.norm(xm - x64) / np.linalg.norm(x64)))
        log(f'config 4 {tag}: {tm * 1e3:.2f} ms ({im["iters"]} it)')
        del e3, fc
    return out


def gnn_application_roofline(app):
    """Roofline of one network application z = M(r) (||r||, the encoder and
    the L message-passing layers: csrc/k_gnn.hip), from its event-timed
    duration. Algorithmic bytes, each array once: ||r|| reads r (8 n); the
    encoder reads r, the vertex features (8 n each) and writes h and Q
    (bfloat16, 2 H n each, H = 32); a layer reads rowptr (4 n), h, Q (2 H n
    each; Q is gathered, counted once), col and the edge features (16 per
    edge) and writes the next h and Q (2 H n each); the last layer writes z
    and reads r and d instead of writing h and Q (24 n)."""
    n, nnz, L, H = app['vertices'], app['edges'], app['layers'], 32
    layer_in = n * (4 + 2 * 2 * H) + nnz * 16
    bytes_ = n * 8 + n * (16 + 2 * 2 * H) + (L - 1) * (layer_in + n * 2 * 2 * H) + \
        layer_in + n * 24
    t = app['us_per_application'] * 1e-6
    return dict(bound='hbm', kernel=f'network application (k_gnn_encode + {L} x k_gnn_layer)',
                achieved=bytes_ / t / 1e9, peak=HBM_PEAK / 1e9, unit='GB/s',
                frac=bytes_ / t / HBM_PEAK, traffic=None, bytes_per_launch=bytes_,
                avg_application_s=t)


def learned_config5(weights, seeds=(1,), reps=1,
                    methods=('jacobi', 'mg', 'mg-learned', 'amg'), skip_cpu=False):
    """Learned-preconditioner flexible CG against Jacobi-PCG on unseen
    variable-coefficient problems on the 2048 x 2048 mesh, all to 1e-8:
    'mg-learned' is flexible CG preconditioned by a multigrid V-cycle whose
    smoother is the trained network, 'mg' the same cycle with damped-Jacobi
    smoothing (the method without the network), 'amg' the same cycle on a
    hierarchy built by smoothed aggregation from the operator alone (the
    form that applies to any mesh)."""
    from easier_amd.numeric.preconditioner import learned as lp
    from easier_amd.workloads import learned as wl
    from easier_amd.workloads import poisson as wp
    pre = lp.load(weights)
    out = []
    for s in seeds:
        res = wl.evaluate_compiled(pre, 2048, s, rtol=1e-8, reps=reps, maxiter=50000,
                                   methods=methods)
        res['network_application'] = wl.time_forward(pre, 2048, s, reps=20)
        res['network_roofline'] = gnn_application_roofline(res['network_application'])
        if not skip_cpu and 'jacobi' in res:
            rate = cpu_pcg_rate(wp.varcoef_system(2048, s)[1])
            res['cpu_jacobi_s_per_iter'] = rate
            res['cpu_jacobi_solve_estimate_s'] = rate * res['jacobi']['iters']
        out.append(res)
    return out


def copy_calibration(mib):
    """dst = src for a float64 array of `mib` MiB, 5 launches: reads and
    writes exactly mib MiB each; prints the event-timed bandwidth."""
    import torch
    from easier_amd.core import ops
    n = mib * (1 << 20) // 8
    src = torch.empty(n, dtype=torch.float64, device='cuda').uniform_()
    dst = torch.empty_like(src)
    ops.copy_(dst, src)
    torch.cuda.synchronize()
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
    stream = torch.cuda.current_stream()
    for i in range(5):
        evs[i].record(stream)
        ops.copy_(dst, src)
    evs[5].record(stream)
    torch.cuda.synchronize()
    ts = [evs[i].elapsed_time(evs[i + 1]) / 1e3 for i in range(5)]
    t = sorted(ts)[2]
    print(json.dumps({'copy_bytes_each_way': n * 8, 'median_s': t,
                      'read_plus_write_GBps': 2 * n * 8 / t / 1e9}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=500)
    ap.add_argument('--warmup', type=int, default=10)
    ap.add_argument('--n', type=int, default=2048)
    ap.add_argument('--dtype', choices=['f64', 'f32'], default='f64')
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--cpu-steps', type=int, default=2)
    ap.add_argument('--skip-cpu', action='store_true')
    ap.add_argument('--skip-poisson', action='store_true')
    ap.add_argument('--skip-fp32', action='store_true',
                    help='skip the float32 SWE side figure')
    ap.add_argument('--no-fuse', action='store_true')
    ap.add_argument('--reorder', choices=['rcm', 'blocks', 'none'], default=None,
                    help="cell renumbering (default: 'blocks' with freeze, else 'rcm')")
    ap.add_argument('--no-freeze', action='store_true',
                    help='compile without freeze=True (geometry read as plain arrays '
                         'instead of through dictionaries)')
    ap.add_argument('--profile', action='store_true',
                    help='only warm-up + timed steps (for rocprofv3 runs)')
    ap.add_argument('--configs', default='3,4,5',
                    help='comma list of extra workloads measured after the headline: '
                         '3 (batched PCG), 4 (mixed precision), 5 (learned FCG); '
                         'empty for none')
    ap.add_argument('--weights', default=os.path.join(REPO, 'easier_amd', 'workloads',
                                                      'gnn_smoother_l3_h32.npz'))
    ap.add_argument('--pcg-profile', action='store_true',
                    help='only the Jacobi-PCG iteration of the 2048 x 2048 variable-'
                         'coefficient system in --dtype, 3 captured chunks of 32 '
                         'iterations after one warm-up chunk (for rocprofv3 runs)')
    ap.add_argument('--config3-only', action='store_true',
                    help='only the batched configuration (64 right-hand sides on 1024^2, '
                         '256 meshes), as in the default run (for rocprofv3 runs)')
    ap.add_argument('--copy-calibration', type=int, default=0, metavar='MIB',
                    help='only copy a float64 array of MIB MiB 5 times (the known-size '
                         'kernel that scales the HBM byte counters)')
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) > 1:
        print('bench.py: this build runs on one GPU; --gpus must be 1 '
              '(no multi-GPU path is built)', file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print('bench.py: no GPU visible', file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.config3_only:
        print(json.dumps(batched_config3()), flush=True)
        return
    if args.pcg_profile:
        print(json.dumps(pcg_profile(args.dtype)), flush=True)
        return
    dtype = torch.float64 if args.dtype == 'f64' else torch.float32
    T = 8 if dtype == torch.float64 else 4
    freeze = not args.no_freeze
    reorder = args.reorder or ('blocks' if freeze else 'rcm')
    mesh, geo, eqn, meta = build_swe(args.n, dtype, args.seed, fuse=not args.no_fuse,
                                     freeze=freeze, reorder=reorder)
    nc, ne, nbc = mesh['cells'].shape[0], mesh['src'].shape[0], mesh['bcells'].shape[0]
    for _ in range(args.warmup):
        eqn()
    torch.cuda.synchronize()
    log(f'warm-up {args.warmup} steps done; timing {args.steps} steps')
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record()
    for _ in range(args.steps):
        eqn()
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_s = start.elapsed_time(stop) / 1e3
    ms_per_step = 1e3 * wall / args.steps
    value = nc * args.steps / wall
    log(f'{ms_per_step:.4f} ms/step (events {1e3 * gpu_s / args.steps:.4f}), '
        f'{value:.3e} cells/s')
    h = eqn.h.collect()
    finite = bool(torch.isfinite(h).all())
    stages, step_bytes = swe_bytes(nc, ne, nbc, T)
    kerns = getattr(eqn.easier_engine, 'kernels', None)
    # compulsory bytes of each generated kernel (its own arrays once each: with
    # freeze, dictionary codes and padded edge records instead of the
    # reference layout's per-edge geometry)
    own = [k.algorithmic_bytes() for k in kerns] if kerns else None
    result = {
        'metric': METRIC, 'value': value, 'unit': 'cells/s', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': ms_per_step,
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': args.dtype, 'data': 'synthetic',
        'config': {
            'workload': f'shallow-water RK4, 2-triangle mesh {args.n}x{args.n} '
                        f'({nc} cells, {ne} directed edges, {nbc} wall faces), '
                        f'{args.dtype}, g=9.81, CFL 0.4 (dt={meta["dt"]:.4e}), '
                        'reflective walls' + (', compiled with freeze=True (constant '
                                              'geometry read through dictionaries), '
                                              f'reorder={reorder}' if freeze else ''),
            'freeze': freeze, 'reorder': reorder,
            'cells': nc, 'edges': ne, 'engine': type(eqn.easier_engine).__name__,
            'step_bytes_algorithmic': step_bytes,
            'step_rate_on_algorithmic_bytes': step_bytes / (wall / args.steps) / HBM_PEAK,
            'step_bytes_kernel_compulsory': sum(own) if own else None,
            'step_hbm_fraction': (sum(own) / (wall / args.steps) / HBM_PEAK) if own else None,
            'edges_per_s': ne * args.steps / wall,
            'state_finite': finite,
        },
    }
    if args.profile:
        print(json.dumps(result), flush=True)
        return
    ktimes = time_kernels(eqn, reps=20)
    if ktimes:
        k = max(range(len(ktimes)), key=lambda i: ktimes[i])
        achieved = stages[k] / ktimes[k]
        result['roofline'] = {
            'bound': 'hbm', 'kernel': f'esr_fused_{k} (RK4 stage {k + 1} of 4)',
            'achieved': achieved / 1e9, 'peak': HBM_PEAK / 1e9, 'unit': 'GB/s',
            'frac': achieved / HBM_PEAK,
            'traffic': committed_traffic(args.n, args.dtype, f'esr_fused_{k}',
                                         '_frozen' if freeze else ''),
            'bytes_per_launch': stages[k], 'avg_launch_s': ktimes[k],
            'all_kernels_s': ktimes,
            'kernel_compulsory_bytes': own[k] if own else None,
            'frac_on_kernel_compulsory_bytes': (own[k] / ktimes[k] / HBM_PEAK) if own else None,
        }
        log(f'dominant kernel esr_fused_{k}: {ktimes[k] * 1e6:.1f} us, '
            f'{achieved / 1e9:.0f} GB/s ({achieved / HBM_PEAK:.1%} of 8 TB/s)')
    result['primitives'] = primitives(eqn, with_cpu=not args.skip_cpu)
    result['primitives']['mesh_order'] = reorder
    if not args.skip_cpu:
        log(f'CPU baseline: {args.cpu_steps} oracle steps on the host')
        result['cpu_baseline'] = cpu_baseline_swe(mesh, geo, meta, dtype, args.cpu_steps)
    if not args.skip_poisson:
        log('Poisson config 1')
        result['poisson_config1'] = poisson_config1(args.skip_cpu)
        log('Poisson config 1, GMRES(20)')
        result['poisson_gmres_config1'] = poisson_gmres_config1(args.skip_cpu)
    if not args.skip_fp32 and args.dtype == 'f64':
        log('SWE float32')
        del eqn
        torch.cuda.empty_cache()
        result['swe_fp32'] = swe_side(args.n, args.seed, torch.float32, freeze)
        if freeze:
            log('SWE float64 without freeze')
            torch.cuda.empty_cache()
            result['swe_fp64_unfrozen'] = swe_side(args.n, args.seed, torch.float64, False,
                                                   steps=40, with_primitives=True)
    extra = {c.strip() for c in args.configs.split(',') if c.strip()}
    if '3' in extra:
        result['batched_config3'] = batched_config3()
    if '4' in extra:
        result['mixed_config4'] = mixed_config4(skip_cpu=args.skip_cpu)
    if '5' in extra:
        log('config 5: Jacobi-PCG, multigrid (geometric, learned, aggregation) flexible CG, 2048^2')
        result['learned_config5'] = learned_config5(args.weights, skip_cpu=args.skip_cpu)
        r5 = result['learned_config5'][0]
        if 'jacobi' in r5 and 'mg-learned' in r5:
            result['poisson_solve_2048'] = {
                'workload': 'variable-coefficient P1 Poisson 2048x2048 (4,198,401 '
                            'vertices), seed 1, FP64, rtol 1e-8, from x = 0',
                'jacobi_pcg_s': r5['jacobi']['seconds'],
                'jacobi_pcg_iters': r5['jacobi']['iters'],
                'learned_mg_fcg_s': r5['mg-learned']['seconds'],
                'learned_mg_fcg_iters': r5['mg-learned']['iters'],
                'learned_speedup_vs_jacobi': r5['mg-learned']['speedup_vs_jacobi'],
                'mg_fcg_s': r5.get('mg', {}).get('seconds'),
                'mg_fcg_iters': r5.get('mg', {}).get('iters'),
                'amg_fcg_s': r5.get('amg', {}).get('seconds'),
                'amg_fcg_iters': r5.get('amg', {}).get('iters'),
            }
    print(json.dumps(result), flush=True)


if __name__ == '__main__':
    main()
