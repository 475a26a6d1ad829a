"""GPU parity of the primitive kernels against the oracle (torch CPU ops of
the reference). Bit-exact for gathers and for Reducer rows up to the
sequential limit; aggregators within a few ulps and run-to-run identical."""
import math

import numpy as np
import pytest
import torch

from easier_amd import _native as nat
from easier_amd.core import compiler, ops
from oracle import ops as oops

pytestmark = pytest.mark.gpu


def _plan(idx: np.ndarray, n: int, dev):
    rowptr, perm = compiler.csr_from_idx(idx, n)
    counts = np.diff(rowptr)
    long_rows = np.nonzero(counts > nat.SEQ_ROW_MAX)[0]
    t = lambda a: torch.from_numpy(a.astype(np.int32)).to(dev)
    return (t(rowptr), t(perm), t(long_rows) if long_rows.size else None)


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32, torch.int64,
                                   torch.int32])
@pytest.mark.parametrize('sub', [(), (3,), (2, 2), (5,)])
def test_gather_bit_exact(gpu, dtype, sub):
    print(f'gather {dtype} {sub}', flush=True)
    g = torch.Generator().manual_seed(0)
    n_src, n_out = 1000, 4099
    src = (torch.randn((n_src,) + sub, generator=g, dtype=torch.float64) * 100).to(dtype)
    idx = torch.randint(0, n_src, (n_out,), generator=g)
    out = ops.gather(src.to(gpu), idx.to(torch.int32).to(gpu))
    assert torch.equal(out.cpu(), oops.selector(src, idx))


def test_gather_empty_and_large(gpu):
    src = torch.randn(10, dtype=torch.float64)
    idx = torch.zeros(0, dtype=torch.int32)
    assert ops.gather(src.to(gpu), idx.to(gpu)).shape == (0,)
    n = 3_000_001
    idx = torch.randint(0, 10, (n,), generator=torch.Generator().manual_seed(1))
    out = ops.gather(src.to(gpu), idx.to(torch.int32).to(gpu))
    assert torch.equal(out.cpu(), src[idx])
    # every tail length of the 16-byte vector path (2 float64 / 4 float32 per vector)
    for dt in (torch.float64, torch.float32):
        s = src.to(dt)
        for n in range(1, 12):
            idx = torch.randint(0, 10, (n,), generator=torch.Generator().manual_seed(n))
            out = ops.gather(s.to(gpu), idx.to(torch.int32).to(gpu))
            assert torch.equal(out.cpu(), s[idx]), (dt, n)


@pytest.mark.parametrize('reduce', ['sum', 'prod', 'mean', 'amax', 'amin'])
@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_reducer_bit_exact(gpu, reduce, dtype):
    print(f'reduce {reduce} {dtype}', flush=True)
    g = torch.Generator().manual_seed(2)
    ne, n = 5000, 1200          # rows of ~4 entries, some empty
    idx = torch.randint(0, n, (ne,), generator=g)
    t = torch.randn((ne, 3), generator=g, dtype=torch.float64).to(dtype)
    if reduce in ('amax', 'amin'):
        t[7, 1] = float('nan')
    rowptr, perm, lr = _plan(idx.numpy(), n, gpu)
    out = ops.csr_reduce(t.to(gpu), rowptr, perm, lr, n, reduce).cpu()
    ref = oops.reducer(t, idx, n, reduce)
    assert torch.equal(torch.nan_to_num(out, nan=7.0), torch.nan_to_num(ref, nan=7.0))


@pytest.mark.parametrize('reduce', ['sum', 'prod', 'mean', 'amax', 'amin'])
@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_reducer_sorted_single_value_rows(gpu, reduce, dtype):
    """Sorted index, one value per edge (the compiled Reducer's form, staged
    through LDS per 64 rows): short rows, empty rows, a 64-row group whose
    entries exceed the staging capacity, rows beyond the sequential limit."""
    print(f'reduce sorted {reduce} {dtype}', flush=True)
    g = torch.Generator().manual_seed(6)
    n = 3001
    counts = torch.randint(0, 7, (n,), generator=g)
    counts[:64] = 0
    counts[200:260] = 40                                 # 2,400 entries in one group
    counts[1000] = 300                                   # beyond the sequential limit
    idx = torch.repeat_interleave(torch.arange(n), counts)
    t = torch.randn(idx.shape[0], generator=g, dtype=torch.float64).to(dtype)
    if reduce in ('amax', 'amin'):
        t[5] = float('nan')
    rowptr = torch.from_numpy(np.concatenate([[0], np.cumsum(counts.numpy())]).astype(np.int32))
    lr = torch.nonzero(counts > nat.SEQ_ROW_MAX).reshape(-1).to(torch.int32)
    out = ops.csr_reduce(t.to(gpu), rowptr.to(gpu), None, lr.to(gpu), n, reduce).cpu()
    ref = oops.reducer(t, idx, n, reduce)
    keep = counts <= nat.SEQ_ROW_MAX
    assert torch.equal(torch.nan_to_num(out[keep], nan=7.0), torch.nan_to_num(ref[keep], nan=7.0))
    if reduce == 'sum':
        assert abs(float(out[1000]) - math.fsum(t[idx == 1000].double().tolist())) <= \
            1e-5 * float(t[idx == 1000].double().abs().sum())


def test_reducer_out_keeps_empty_rows(gpu):
    # tests/core/test_module.py:71-86
    idx = torch.tensor([1, 1, 2, 2, 2])
    edge = torch.randn((5, 3, 4), generator=torch.Generator().manual_seed(3))
    out = torch.rand((3, 3, 4), generator=torch.Generator().manual_seed(4))
    ref = oops.reducer(edge, idx, 3, out=out.clone())
    rowptr, perm, lr = _plan(idx.numpy(), 3, gpu)
    o = out.to(gpu)
    res = ops.csr_reduce(edge.to(gpu), rowptr, perm, lr, 3, 'sum', out=o,
                         keep_empty=True)
    assert res.data_ptr() == o.data_ptr()
    assert torch.equal(res.cpu(), ref)
    # without out=, empty rows are zeros
    res = ops.csr_reduce(edge.to(gpu), rowptr, perm, lr, 3, 'sum').cpu()
    assert torch.equal(res, oops.reducer(edge, idx, 3))


def test_reducer_long_rows(gpu):
    g = torch.Generator().manual_seed(5)
    ne = 100_000
    idx = torch.randint(0, 4, (ne,), generator=g)        # 4 rows of ~25k
    idx[:10] = 5                                          # one short row
    t = torch.randn(ne, generator=g, dtype=torch.float64)
    rowptr, perm, lr = _plan(idx.numpy(), 6, gpu)
    assert lr is not None and lr.shape[0] == 4
    out = ops.csr_reduce(t.to(gpu), rowptr, perm, lr, 6, 'sum').cpu()
    ref = oops.reducer(t, idx, 6)
    assert torch.equal(out[5], ref[5]) and out[4] == 0
    for r in range(4):
        exact = math.fsum(t[idx == r].tolist())
        assert abs(float(out[r]) - exact) <= 1e-13 * float(t[idx == r].abs().sum())
    out2 = ops.csr_reduce(t.to(gpu), rowptr, perm, lr, 6, 'sum').cpu()
    assert torch.equal(out, out2)                         # deterministic


@pytest.mark.parametrize('op', ['sum', 'norm', 'max', 'min', 'prod'])
@pytest.mark.parametrize('n', [1, 777, 300_001])
def test_aggregators(gpu, op, n):
    g = torch.Generator().manual_seed(6)
    t = torch.randn((n, 2), generator=g, dtype=torch.float64)
    if op == 'prod':
        t = 1 + 1e-4 * t
    a = ops.aggregate(op, t.to(gpu)).cpu()
    b = oops.aggregate(op, t)
    assert a.shape == b.shape == (1, 2)
    if op in ('max', 'min'):
        assert torch.equal(a, b)
    else:
        torch.testing.assert_close(a, b, rtol=1e-12, atol=0)
    a2 = ops.aggregate(op, t.to(gpu)).cpu()
    assert torch.equal(a, a2)


@pytest.mark.parametrize('op', ['sum', 'norm', 'max', 'min'])
@pytest.mark.parametrize('n', [1, 3, 777, 300_001])
@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_aggregators_one_column(gpu, op, n, dtype):
    """One column, float64 and float32, aligned and offset views, lengths
    that leave every remainder of a 16-byte vector."""
    g = torch.Generator().manual_seed(8)
    t = torch.randn(n + 1, generator=g, dtype=torch.float64).to(dtype)
    tol = 1e-12 if dtype == torch.float64 else 1e-4
    dev = t.to(gpu)
    for host, view in ((t[:n], dev[:n]), (t[1:], dev[1:])):
        a = ops.aggregate(op, view.reshape(-1, 1)).cpu().reshape(-1)
        b = oops.aggregate(op, host.reshape(-1, 1).double()).reshape(-1)
        if op in ('max', 'min'):
            assert float(a[0]) == float(b[0])
        else:
            assert abs(float(a[0]) - float(b[0])) <= tol * max(abs(float(b[0])), 1.0)
        assert torch.equal(a, ops.aggregate(op, view.reshape(-1, 1)).cpu().reshape(-1))


@pytest.mark.parametrize('op', ['sum', 'max', 'min'])
@pytest.mark.parametrize('dtype', [torch.int64, torch.int32])
def test_aggregators_one_column_integers(gpu, op, dtype):
    """Integer columns long enough for the vector-load stage, aligned and
    offset by one element: exact."""
    g = torch.Generator().manual_seed(9)
    t = torch.randint(-1000, 1000, (300_002,), generator=g, dtype=torch.int64).to(dtype)
    dev = t.to(gpu)
    for host, view in ((t[:-1], dev[:-1]), (t[1:], dev[1:])):
        a = ops.aggregate(op, view.reshape(-1, 1)).cpu().reshape(-1)
        b = oops.aggregate(op, host.reshape(-1, 1)).reshape(-1)
        assert int(a[0]) == int(b[0]), (int(a[0]), int(b[0]))


@pytest.mark.parametrize('op', ['add', 'sub', 'mul', 'div'])
@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_ewise_arithmetic_bit_exact(gpu, op, dtype):
    g = torch.Generator().manual_seed(7)
    a = torch.randn((1000, 3), generator=g, dtype=torch.float64).to(dtype)
    b = torch.randn((1000, 1), generator=g, dtype=torch.float64).to(dtype)
    f = {'add': torch.add, 'sub': torch.sub, 'mul': torch.mul, 'div': torch.div}[op]
    assert torch.equal(ops.ewise(op, a.to(gpu), b.to(gpu)).cpu(), f(a, b))
    assert torch.equal(ops.ewise(op, a.to(gpu), 0.37).cpu(), f(a, 0.37))
    assert torch.equal(ops.ewise(op, 2.5, a.to(gpu)).cpu(), f(torch.tensor(2.5, dtype=dtype), a))
    s = torch.tensor([1.7], dtype=dtype)
    assert torch.equal(ops.ewise(op, a.to(gpu), s.to(gpu)).cpu(), f(a, s))


def test_ewise_misc(gpu):
    g = torch.Generator().manual_seed(8)
    a = torch.randn(4000, generator=g, dtype=torch.float64)
    ad = a.to(gpu)
    # sqrt is correctly rounded (IEEE 754), checked bit for bit against
    # numpy's; torch's vectorised CPU sqrt (the reference CPU backend) is not
    # correctly rounded on AVX-512 hosts and may differ by one ulp.
    sq = ops.ewise('sqrt', ops.ewise('abs', ad)).cpu()
    assert np.array_equal(sq.numpy(), np.sqrt(np.abs(a.numpy())))
    ulps = (sq.view(torch.int64) - a.abs().sqrt().view(torch.int64)).abs()
    assert int(ulps.max()) <= 1
    assert torch.equal(ops.ewise('pow', ad, 2).cpu(), a ** 2)
    assert torch.equal(ops.ewise('sign', ad).cpu(), torch.sign(a))
    assert torch.equal(ops.ewise('neg', ad).cpu(), -a)
    torch.testing.assert_close(ops.ewise('exp', ad).cpu(), torch.exp(a), rtol=4e-16, atol=0)
    c = ops.ewise('lt', ad, 0.0)
    assert c.dtype == torch.bool and torch.equal(c.cpu(), a < 0)
    w = ops.ewise('where', c, ad, 0.0).cpu()
    assert torch.equal(w, torch.where(a < 0, a, 0.0))
    # strided view in and out
    M = torch.randn((50, 4), generator=g, dtype=torch.float64)
    Md = M.to(gpu)
    col = ops.ewise('mul', Md[:, 1], Md[:, 3]).cpu()
    assert torch.equal(col, M[:, 1] * M[:, 3])
    ops.copy_(Md[:, 0], Md[:, 2])
    M[:, 0] = M[:, 2]
    assert torch.equal(Md.cpu(), M)
    # dtype promotion: f32 tensor op f64 tensor -> f64
    x32 = torch.randn(10, generator=g)
    r = ops.ewise('add', x32.to(gpu), a[:10].to(gpu))
    assert r.dtype == torch.float64 and torch.equal(r.cpu(), x32 + a[:10])
    # casts
    i = ops.ewise('copy', ad, out=torch.empty(4000, dtype=torch.int64, device=gpu))
    assert torch.equal(i.cpu(), a.to(torch.int64))


def test_gemv_rows(gpu):
    g = torch.Generator().manual_seed(9)
    V = torch.randn((1000, 20), generator=g, dtype=torch.float64)
    y = torch.randn((20, 1), generator=g, dtype=torch.float64)
    out = ops.gemv_rows(V[:, :7].to(gpu), y[:7].to(gpu)).cpu()
    torch.testing.assert_close(out, (V[:, :7] @ y[:7]).squeeze(-1), rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_primitives_at_full_size(gpu, dtype):
    """The benchmark's primitives at their size, bit for bit against the
    oracle: the cell -> edge gather of the 2048 x 2048 2-triangle mesh's
    25,157,632 directed dual edges (16-byte vector path) and the edge -> cell
    sum over the edges sorted by destination (LDS-staged path), one value per
    edge, 8,388,608 rows."""
    from easier_amd.workloads import mesh as wmesh
    print(f'primitives full size {dtype}', flush=True)
    mesh = wmesh.tri2(2048)
    src = torch.as_tensor(mesh['src']).long()
    dst = torch.as_tensor(mesh['dst']).long()
    nc = mesh['cells'].shape[0]
    order = torch.sort(dst, stable=True).indices   # the compiled edge order
    src, dst = src[order], dst[order]
    g = torch.Generator().manual_seed(11)
    cells = torch.randn(nc, generator=g, dtype=torch.float64).to(dtype)
    got = ops.gather(cells.to(gpu), src.to(torch.int32).to(gpu)).cpu()
    want = oops.selector(cells, src)
    assert torch.equal(got, want)
    counts = torch.bincount(dst, minlength=nc)
    rowptr = torch.cat([torch.zeros(1, dtype=torch.long), torch.cumsum(counts, 0)]).to(torch.int32)
    lr = torch.zeros(0, dtype=torch.int32)
    out = ops.csr_reduce(got.to(gpu), rowptr.to(gpu), None, lr.to(gpu), nc, 'sum').cpu()
    ref = oops.reducer(want, dst, nc)
    assert torch.equal(out, ref)


def test_ewise_integer_pow(gpu):
    """Integer powers follow torch's integer arithmetic: a negative exponent
    gives 1 for a = 1, +-1 for a = -1 and 0 otherwise; large exponents wrap
    modulo 2^64; a negative scalar exponent is an error."""
    a = torch.tensor([2, 1, -1, 0, 3, -2, -1, 7, 3, 5], dtype=torch.int64)
    b = torch.tensor([-1, -2, -3, -1, 0, 63, 64, 3, 2 ** 40, 2 ** 62 + 1], dtype=torch.int64)
    got = ops.ewise('pow', a.to(gpu), b.to(gpu)).cpu()
    assert torch.equal(got, torch.pow(a, b)), (got, torch.pow(a, b))
    assert int(got[8]) == -7860764868738023423
    a32 = torch.arange(-20, 20, dtype=torch.int32)
    for e in (0, 1, 2, 3, 7, 31):
        assert torch.equal(ops.ewise('pow', a32.to(gpu), e).cpu(), a32 ** e), e
    with pytest.raises(RuntimeError, match='negative integer powers'):
        ops.ewise('pow', a.to(gpu), -2)
