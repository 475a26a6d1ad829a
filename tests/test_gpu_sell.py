"""The SELL-64 form of the Linsys operator (csrc/esr_krylov.h: sell_row,
csrc/k_sell.hip) and the PCG iteration on it.

* Layout: soff, cnt, colv and val read back from the device equal a host
  construction of the format's definition (slice s = rows 64s..64s+63,
  width = its longest row, entry k of row 64s+l at soff[s] + 64k + l, padding
  value 0), with int16 offsets and with int32 indices, on ragged rows
  (empty rows, rows longer than the 8 unrolled slots, a partial last slice).
* The fused Jacobi-PCG on the SELL form gives bit for bit the iterates,
  residuals and step counts of the CSR form (which the other solver tests pin
  to the oracle): the config-1 system (two-kernel iteration), a 160,801-vertex
  system (three-kernel iteration), a randomly renumbered system whose column
  offsets do not fit int16, float64 and float32; after the operator's values
  are changed in place between solves; and the mixed-precision CG's float32
  inner iteration.
"""
import numpy as np
import pytest
import torch

import easier_amd as esr
from easier_amd.numeric.solver import CG
from easier_amd.numeric.solver import fused as F
from easier_amd.numeric.solver.mixed import MixedPrecisionCG
from easier_amd.workloads import mesh as wmesh
from easier_amd.workloads import poisson as wp

pytestmark = pytest.mark.gpu


def _host_sell(rowptr, col, Ae, idx16):
    n = rowptr.shape[0] - 1
    ns = (n + 63) // 64
    lens = np.diff(rowptr)
    lens_p = np.zeros(ns * 64, np.int64)
    lens_p[:n] = lens
    width = lens_p.reshape(ns, 64).max(axis=1)
    soff = np.zeros(ns + 1, np.int64)
    soff[1:] = np.cumsum(64 * width)
    m = max(int(soff[-1]), 1)
    colv = np.zeros(m, np.int64)
    val = np.zeros(m, Ae.dtype)
    for r in range(ns * 64):
        s, l = divmod(r, 64)
        pad = r if r < n else n - 1
        for k in range(int(width[s])):
            e = soff[s] + 64 * k + l
            if r < n and k < lens[r]:
                c = col[rowptr[r] + k]
                val[e] = Ae[rowptr[r] + k]
            else:
                c = pad
            colv[e] = c - r if idx16 else c
    return soff, lens.astype(np.uint8), colv, val


class _CSR:
    def __init__(self, rowptr, col, Ae, dev):
        self.rowptr = torch.from_numpy(rowptr.astype(np.int32)).to(dev)
        self.col = torch.from_numpy(col.astype(np.int32)).to(dev)
        self.Ae = torch.from_numpy(Ae).to(dev)
        self.Av = torch.ones(rowptr.shape[0] - 1, dtype=self.Ae.dtype, device=dev)
        self.n = rowptr.shape[0] - 1


@pytest.mark.parametrize('spread', [300, 100000])
@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_sell_layout_matches_its_definition(gpu, spread, dtype):
    rng = np.random.default_rng(7)
    n = 1000
    lens = rng.integers(0, 13, n)
    lens[5:70] = 0                       # a whole slice of empty rows
    lens[300] = 40                       # a row far beyond the unrolled slots
    rowptr = np.zeros(n + 1, np.int64)
    rowptr[1:] = np.cumsum(lens)
    rows = np.repeat(np.arange(n), lens)
    col = np.clip(rows + rng.integers(-spread, spread + 1, rows.shape[0]), 0, n - 1)
    if spread > 30000:                   # offsets beyond int16 need more rows
        n2 = 70000
        col = rng.integers(0, n2, rows.shape[0])
        rowptr = np.concatenate([rowptr, np.full(n2 - n, rowptr[-1])])
    Ae = rng.standard_normal(rows.shape[0]).astype(dtype)
    print(f'SELL layout: n={rowptr.shape[0] - 1} nnz={rows.shape[0]} spread={spread} '
          f'{np.dtype(dtype).name}', flush=True)
    csr = _CSR(rowptr, col, Ae, 'cuda')
    sell = F.LinsysSELL(csr)
    assert sell.idx16 == (1 if spread < 30000 else 0)
    soff, cnt, colv, val = _host_sell(rowptr, col, Ae, sell.idx16)
    assert sell.slots == int(soff[-1])
    np.testing.assert_array_equal(sell.soff.cpu().numpy(), soff)
    np.testing.assert_array_equal(sell.cnt.cpu().numpy(), cnt)
    np.testing.assert_array_equal(sell.colv.cpu().numpy().astype(np.int64), colv)
    np.testing.assert_array_equal(sell.val.cpu().numpy(), val)
    # new values through esr_sell_values: the same layout of the new values
    Ae2 = (Ae * 3 + 1).astype(dtype)
    sell.refresh(torch.from_numpy(Ae2).cuda())
    np.testing.assert_array_equal(sell.val.cpu().numpy(),
                                  _host_sell(rowptr, col, Ae2, sell.idx16)[3])


def _permuted(sys_, seed):
    n = sys_['Av'].shape[0]
    perm = np.random.default_rng(seed).permutation(n)
    row, col = perm[sys_['row']], perm[sys_['col']]
    order = np.lexsort((col, row))
    out = dict(row=row[order], col=col[order], Ae=sys_['Ae'][order])
    for k in ('Av', 'b'):
        v = np.empty_like(sys_[k])
        v[perm] = sys_[k]
        out[k] = v
    return out


def _solve(sys_, dtype, use_sell, reorder, maxiter, rtol=1e-8, rescale=None):
    old = F.FusedCG.use_sell
    F.FusedCG.use_sell = use_sell
    try:
        eqn = wp.PoissonP1(sys_, dtype=dtype)
        sol = CG(eqn.A, eqn.b, eqn.x, eqn.M)
        esr.compile([eqn, sol], 'hip', reorder=reorder)
    finally:
        F.FusedCG.use_sell = old
    assert sol.easier_fused is not None
    info = sol.solve(rtol=rtol, maxiter=maxiter)
    assert (sol.easier_fused.sell is not None) == use_sell
    out = [(info['iters'], info['residual'], eqn.x.collect().clone())]
    if rescale is not None:
        eqn.Ae.data.mul_(rescale)
        eqn.Av.data.mul_(rescale)
        eqn.x.data.zero_()
        info = sol.solve(rtol=rtol, maxiter=maxiter)
        out.append((info['iters'], info['residual'], eqn.x.collect().clone()))
    if use_sell:
        idx16 = sol.easier_fused.sell.idx16
    else:
        idx16 = None
    return out, idx16


CASES = {
    'config1': dict(n=100, perm=False, reorder='rcm', maxiter=5000, idx16=1),
    'large': dict(n=400, perm=False, reorder='rcm', maxiter=300, idx16=1),
    'int32': dict(n=200, perm=True, reorder='none', maxiter=400, idx16=0),
}


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
@pytest.mark.parametrize('case', list(CASES))
def test_sell_pcg_is_bit_identical_to_csr(gpu, case, dtype):
    c = CASES[case]
    _, sys_ = wp.config1_system(c['n'], seed=0)
    if c['perm']:
        sys_ = _permuted(sys_, 3)
    print(f'SELL vs CSR PCG: {case} ({sys_["Av"].shape[0]} rows), {dtype}', flush=True)
    a, idx16 = _solve(sys_, dtype, True, c['reorder'], c['maxiter'], rescale=2.5)
    b, _ = _solve(sys_, dtype, False, c['reorder'], c['maxiter'], rescale=2.5)
    assert idx16 == c['idx16']
    for (ia, ra, xa), (ib, rb, xb) in zip(a, b):
        print(f'  iters {ia} / {ib}, residual {ra:.6e} / {rb:.6e}', flush=True)
        assert ia == ib and ra == rb
        assert torch.equal(xa, xb)
    # the rescaled operator is another system: the solve saw the new values
    assert not torch.equal(a[0][2], a[1][2])


def test_sell_mixed_precision_inner_is_bit_identical_to_csr(gpu):
    _, sys_, _ = wp.varcoef_system(64, seed=2)
    print('SELL vs CSR mixed-precision CG, 64x64 variable coefficient', flush=True)
    res = []
    for use_sell in (True, False):
        old = F.FusedCG.use_sell
        F.FusedCG.use_sell = use_sell
        try:
            eqn = wp.PoissonP1(sys_)
            sol = MixedPrecisionCG(eqn.A, eqn.b, eqn.x, eqn.M)
            esr.compile([eqn, sol], 'hip')
        finally:
            F.FusedCG.use_sell = old
        info = sol.solve(rtol=1e-10)
        assert (sol.easier_state['sell'] is not None) == use_sell
        res.append((info['iters'], info['outer'], info['residual'], eqn.x.collect()))
    print(f'  {res[0][:3]} / {res[1][:3]}', flush=True)
    assert res[0][:3] == res[1][:3]
    assert torch.equal(res[0][3], res[1][3])
