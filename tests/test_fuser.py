"""The fusion compiler on the host: which forwards it takes, how it stages
them, and that the generated kernels compile for gfx950 (hipRTC, no device).
"""
import ctypes

import numpy as np
import pytest
import torch

import easier_amd as esr
from easier_amd import _native as nat
from easier_amd.core import fuser
from easier_amd.core.compiler import compile_program
from easier_amd.numeric import Linsys
from easier_amd.numeric.solver import CG
from easier_amd.workloads import mesh as wmesh
from easier_amd.workloads import poisson as wp
from easier_amd.workloads import swe as wswe
from oracle import swe as oswe


def _rtc_ok(src):
    nb = ctypes.c_int64()
    rc = nat.lib().esr_rtc_check(src.encode(), fuser.rtc_options('gfx950').encode(),
                                 ctypes.byref(nb))
    assert rc == 0, nat.lib().esr_last_error().decode()
    assert nb.value > 0


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
@pytest.mark.parametrize('kind', ['tri2', 'tri4'])
def test_swe_step_fuses_into_four_kernels(dtype, kind):
    mesh = getattr(wmesh, kind)(5)
    geo = oswe.geometry(mesh)
    eqn = wswe.ShallowWaterEquation(mesh, geo, dt=0.001, g=9.81, dtype=dtype)
    prog = compile_program([eqn], 'rcm', torch.device('cpu'))
    p, low, src = fuser.lower_module(prog.modules[0], prog.graphs[0], prog)
    # one kernel per RK4 stage: the stage inputs are gathered by the next one
    assert [k.stage for k in low.kernels] == [0, 1, 2, 3]
    # written per field: the stage deltas d1, d2, the partial sum
    # (d1 + d2) + d3 and the last stage input U3 = h + dt d3; U1 and U2 are
    # recomputed by the gathers of the next stage from h and d1 / d2, and
    # U3 is not, as the last kernel reassigns h
    assert sum(v.mat for v in p.values) == 12
    assert len(low.remat) == 6
    nc = prog.space_len[low.kernels[0].space]
    cell_arrays = [sum(len(k.access[n]) for n, _, t in k.params if t.numel() == nc)
                   for k in low.kernels]
    assert cell_arrays == [7, 10, 16, 13]
    assert [st.stage for st in p.stores] == [3, 3, 3]
    ctype = 'double' if dtype == torch.float64 else 'float'
    assert f'const {ctype}' in src
    _rtc_ok(src)


@pytest.mark.parametrize('nb,G', [(1, 16), (127, 4), (4096, 16), (4100, 64), (32768, 256)])
def test_group_remap_is_a_permutation(nb, G):
    # esr_group_remap of the generated code, evaluated for every workgroup
    full = (nb // (8 * G)) * 8 * G
    tiles = [b if b >= full else ((b // 8 // G) * 8 + b % 8) * G + (b // 8) % G
             for b in range(nb)]
    assert sorted(tiles) == list(range(nb))
    assert f'esr_group_remap(i64 b, i64 nb, i64 G)' in fuser._CODE_HEADER


def test_sparse_row_bitmap_finds_every_row():
    class _Red:
        pass
    rng = np.random.default_rng(5)
    n = 5000
    counts = np.where(rng.random(n) < 0.05, rng.integers(1, 4, n), 0)
    red = _Red()
    red.easier_host_rowptr = np.concatenate([[0], np.cumsum(counts)])
    red.easier_rowptr = torch.zeros(1, dtype=torch.int32)
    bm, pre, crow = (t.numpy() for t in fuser._sparse_rows(red))
    words = bm.view(np.uint64)
    for r in range(n):
        w, b = r >> 6, r & 63
        present = bool((int(words[w]) >> b) & 1)
        assert present == (counts[r] > 0)
        if present:      # what the generated kernel computes
            c = int(pre[w]) + bin(int(words[w]) & ((1 << b) - 1)).count('1')
            assert (crow[c], crow[c + 1]) == (red.easier_host_rowptr[r],
                                              red.easier_host_rowptr[r + 1])


@pytest.mark.parametrize('opts', [
    dict(stage_edges=False, nt_loads=False, nt_stores=False, block=256, xcd_remap=True,
         xcd_group=0, rows_per_thread=1),
    dict(stage_edges=True, nt_loads=False, nt_stores=True, block=128, xcd_remap=False,
         xcd_group=0, rows_per_thread=1),
    dict(stage_edges=True, nt_loads=True, nt_stores=False, block=512, xcd_remap=True,
         xcd_group=0, rows_per_thread=2),
    dict(stage_edges=False, nt_loads=False, nt_stores=True, block=256, xcd_remap=False,
         xcd_group=64, rows_per_thread=1),
    dict(stage_edges=False, nt_loads=False, nt_stores=True, block=256, xcd_remap=False,
         xcd_group=64, rows_per_thread=2),
    dict(stage_edges=False, nt_loads=False, nt_stores=False, block=128, xcd_remap=False,
         xcd_group=0, rows_per_thread=2),
    dict(stage_edges=False, nt_loads=False, nt_stores=True, block=256, xcd_remap=False,
         xcd_group=64, rows_per_thread=1, pack_intermediates=True)])
def test_code_generation_options(opts):
    mesh = wmesh.tri2(7)
    geo = oswe.geometry(mesh)
    eqn = wswe.ShallowWaterEquation(mesh, geo, dt=0.001, g=9.81)
    prog = compile_program([eqn], 'rcm', torch.device('cpu'))
    saved = dict(fuser.OPTIONS)
    try:
        fuser.OPTIONS.update(opts)
        p, low, src = fuser.lower_module(prog.modules[0], prog.graphs[0], prog)
    finally:
        fuser.OPTIONS.update(saved)
    ksrc = ''.join(k.src for k in low.kernels)
    assert all(k.block == opts['block'] for k in low.kernels)
    assert all(k.staged == opts['stage_edges'] for k in low.kernels)
    # two rows per thread unless the edges are staged through LDS
    rpt = 2 if opts['rows_per_thread'] == 2 and not opts['stage_edges'] else 1
    assert all(k.rpt == rpt for k in low.kernels)
    assert ('esr_ld2(' in ksrc) == (rpt == 2)
    assert ('__syncthreads' in ksrc) == opts['stage_edges']
    assert ('__builtin_nontemporal_store' in ksrc or 'esr_st2nt(' in ksrc) == opts['nt_stores']
    assert ('esr_remap(blockIdx.x' in src) == (opts['xcd_remap'] and not opts['xcd_group'])
    assert ('esr_group_remap(blockIdx.x' in src) == bool(opts['xcd_group'])
    # the staged edge range of one block fits its LDS arrays: 3 edges per row
    for k in low.kernels:
        assert k.lds_bytes <= fuser._LDS_BUDGET
    _rtc_ok(src)


def test_stores_wait_for_gathers_of_old_contents():
    class M(esr.Module):
        def __init__(self):
            super().__init__()
            idx = torch.tensor([1, 0, 2, 2, 0, 1])
            self.red = esr.Reducer(torch.tensor([0, 0, 1, 1, 2, 2]), 3)
            self.sel = esr.Selector(idx)
            self.u = esr.Tensor(torch.rand(3, dtype=torch.double), mode='partition')
            self.w = esr.Tensor(torch.rand(6, dtype=torch.double), mode='partition')

        def forward(self):
            s = self.red(self.sel(self.u) * self.w)
            self.u[:] = self.u + s          # reads u at other rows first
            t = self.red(self.sel(self.u) * self.w)
            self.u[:] = self.u * 2 + t

    m = M()
    prog = compile_program([m], 'rcm', torch.device('cpu'))
    p, low, src = fuser.lower_module(prog.modules[0], prog.graphs[0], prog)
    assert [st.stage for st in p.stores] == [1, 1]
    assert len(low.kernels) == 2
    _rtc_ok(src)


def test_unsupported_forwards_keep_the_node_engine():
    mesh, sys_ = wp.config1_system(6)
    eqn = wp.PoissonP1(sys_)
    sol = CG(eqn.A, eqn.b, eqn.x, eqn.M)
    prog = compile_program([eqn, sol], 'rcm', torch.device('cpu'))
    for m, g in zip(prog.modules, prog.graphs):
        name = m.easier_hint_name
        if name.endswith('.step') or name.endswith('.init') or name.endswith('.update'):
            with pytest.raises(fuser.Unsupported):
                fuser.lower_module(m, g, prog)      # aggregators


def test_linsys_apply_fuses():
    mesh, sys_ = wp.config1_system(6)

    class Apply(esr.Module):
        def __init__(self):
            super().__init__()
            self.p = wp.PoissonP1(sys_)
            self.y = esr.Tensor(esr.zeros((sys_['Av'].shape[0],), dtype=torch.double),
                                mode='partition')

        def forward(self):
            self.y[:] = self.p.A(self.p.b)

    m = Apply()
    prog = compile_program([m], 'rcm', torch.device('cpu'))
    idx = [i for i, mm in enumerate(prog.modules) if mm is m][0]
    p, low, src = fuser.lower_module(prog.modules[idx], prog.graphs[idx], prog)
    assert len(low.kernels) == 1
    _rtc_ok(src)


@pytest.mark.parametrize('remat,late_eval,mat,arrays', [
    (False, False, 18, [10, 16, 16, 13]), (True, False, 15, [7, 13, 16, 13]),
    (False, True, 18, [10, 16, 16, 13]), (True, True, 12, [7, 10, 16, 13])])
def test_recomputed_gathers_and_late_evaluation(remat, late_eval, mat, arrays):
    """Each option alone and together on the SWE step: values written and
    cell arrays read or written per kernel (the algorithmic bytes only fall)."""
    mesh = wmesh.tri2(6)
    geo = oswe.geometry(mesh)
    eqn = wswe.ShallowWaterEquation(mesh, geo, dt=0.001, g=9.81)
    prog = compile_program([eqn], 'rcm', torch.device('cpu'))
    saved = dict(fuser.OPTIONS)
    try:
        fuser.OPTIONS.update(remat=remat, late_eval=late_eval)
        p, low, src = fuser.lower_module(prog.modules[0], prog.graphs[0], prog)
    finally:
        fuser.OPTIONS.update(saved)
    nc = prog.space_len[low.kernels[0].space]
    assert sum(v.mat for v in p.values) == mat
    assert [sum(len(k.access[n]) for n, _, t in k.params if t.numel() == nc)
            for k in low.kernels] == arrays
    _rtc_ok(src)


def test_recompute_refused_when_the_kernel_reassigns_the_source():
    """A gathered value over a tensor that the gathering kernel reassigns is
    written to a buffer (its gathers read the old contents at other rows);
    over tensors left alone, the gathers recompute it."""
    n = 40

    class Two(esr.Module):
        def __init__(self):
            super().__init__()
            g = torch.Generator().manual_seed(3)
            self.sel = esr.Selector(torch.randint(0, n, (3 * n,), generator=g))
            self.red = esr.Reducer(torch.arange(3 * n) % n, n)
            self.u = esr.Tensor(torch.rand(n, dtype=torch.float64, generator=g),
                                mode='partition')
            self.v = esr.Tensor(torch.rand(n, dtype=torch.float64, generator=g),
                                mode='partition')
            self.w = esr.Tensor(torch.rand(n, dtype=torch.float64, generator=g),
                                mode='partition')
            self.x = esr.Tensor(torch.zeros(n, dtype=torch.float64), mode='partition')

        def forward(self):
            self.x[:] = self.red(self.sel(self.w * 2 + self.v))
            self.u[:] = self.red(self.sel(self.u * 3))

    m = Two()
    prog = compile_program([m], 'rcm', torch.device('cpu'))
    p, low, src = fuser.lower_module(prog.modules[0], prog.graphs[0], prog)
    gathered = [u.args[0] for u in p.values if u.kind == 'gather']
    a, b = gathered
    assert a.id in low.remat and not a.mat
    assert b.id not in low.remat and b.mat
    _rtc_ok(src)


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_frozen_constants_are_dictionary_encoded_bit_exactly(dtype):
    """compile(..., freeze=True): the SWE geometry (alpha, sx, sy per edge,
    area per cell) is read through dictionaries
    whose tables reproduce every value bit for bit; the state (written by
    the forward) is not encoded; the kernels compile and move fewer bytes."""
    mesh = wmesh.tri2(64)
    geo = oswe.geometry(mesh)
    eqn = wswe.ShallowWaterEquation(mesh, geo, dt=0.001, g=9.81, dtype=dtype)
    prog = compile_program([eqn], 'rcm', torch.device('cpu'))
    _, low0, _ = fuser.lower_module(prog.modules[0], prog.graphs[0], prog)
    plain = [k.algorithmic_bytes() for k in low0.kernels]
    prog.freeze = True
    _, low, src = fuser.lower_module(prog.modules[0], prog.graphs[0], prog)
    frozen = [k.algorithmic_bytes() for k in low.kernels]
    names = {id(getattr(eqn, k)): k for k in ('alpha', 'sx', 'sy', 'bsx', 'bsy', 'area',
                                                'h', 'uh', 'vh')}
    encoded = {names[i] for i in low.enc}
    # the 256 wall faces are under the size at which a dictionary pays
    assert encoded == {'alpha', 'sx', 'sy', 'area'}
    bits = torch.int64 if dtype == torch.float64 else torch.int32
    for i, enc in low.enc.items():
        t = [getattr(eqn, k) for k in names.values() if id(getattr(eqn, k)) == i][0].data
        got = enc.lut[0].expand(t.shape) if enc.codes is None else enc.lut[enc.codes.long()]
        assert torch.equal(got.view(bits), t.view(bits)), names[i]
    # the edge arrays with several values share one code array (in float32
    # alpha rounds to 1/2 everywhere and is a constant); area holds one value
    codes = {id(low.enc[id(getattr(eqn, k))].codes) for k in ('alpha', 'sx', 'sy')
             if low.enc[id(getattr(eqn, k))].codes is not None}
    assert len(codes) == 1
    assert low.enc[id(eqn.area)].codes is None
    assert all(f < p for f, p in zip(frozen, plain)), (frozen, plain)
    assert '__shared__' in src and '__syncthreads();' in src
    _rtc_ok(src)


def test_frozen_encoding_skips_written_and_many_valued_tensors():
    """A tensor the forward assigns is not a constant; an edge array with
    more distinct values than the dictionary holds is read as an array."""
    mesh = wmesh.tri2(40, jitter=0.2, seed=1)
    geo = oswe.geometry(mesh)
    eqn = wswe.ShallowWaterEquation(mesh, geo, dt=0.001, g=9.81, dtype=torch.float64)
    prog = compile_program([eqn], 'rcm', torch.device('cpu'))
    prog.freeze = True
    _, low, src = fuser.lower_module(prog.modules[0], prog.graphs[0], prog)
    enc_ids = set(low.enc)
    for k in ('h', 'uh', 'vh', 'alpha', 'sx', 'sy', 'area'):
        assert id(getattr(eqn, k)) not in enc_ids, k      # jittered: all distinct
    assert np.unique(geo['alpha'].numpy()).size > 256
    written = fuser._written_tensors(prog)
    assert {id(eqn.h), id(eqn.uh), id(eqn.vh)} <= written
    assert id(eqn.sx) not in written
    _rtc_ok(src)


def test_halo_layout_indexes_every_source_row():
    """reorder='blocks' + freeze: every edge slot's LDS index points at its
    source row -- the tile's own rows first, then the tile's halo list."""
    mesh = wmesh.tri2(50)
    geo = oswe.geometry(mesh)
    eqn = wswe.ShallowWaterEquation(mesh, geo, dt=0.001, g=9.81, dtype=torch.float64)
    prog = compile_program([eqn], 'blocks', torch.device('cpu'))
    sel, red = eqn.gather_src, eqn.scatter
    B, S = 256, 4
    lay = fuser._halo_layout(prog, sel, red, S, B, 256, torch.device('cpu'))
    assert lay is not None
    loc, halo, hs = (x.numpy() if hasattr(x, 'numpy') else x for x in lay)
    slot = fuser._ell_slots(prog, red, S)
    n = slot.shape[0]
    idx = np.asarray(sel.easier_host_idx)
    loc = loc.reshape(n, S).astype(np.int64)
    halo = halo.reshape(-1, hs)
    for r in range(n):
        t = r // B
        for k in range(S):
            e = slot[r, k]
            if e < 0:
                continue
            i = loc[r, k]
            got = t * B + i if i < B else halo[t, i - B]
            assert got == idx[e], (r, k)
    inside = np.mean(loc[slot >= 0] < B)
    assert inside > 0.7
