"""esr_aggregate_noalloc (the aggregator entry point that allocates nothing)
against esr_aggregate with a caller workspace, through the C ABI: two calls
in a row on the library's scratch array, the second on a smaller input, must
give esr_aggregate's bits; requests beyond the scratch size are rejected on
the host before any launch."""
import ctypes

import pytest
import torch

from easier_amd import _native as nat

pytestmark = pytest.mark.gpu


def _with_workspace(op, x, p, stream):
    n, width = x.shape
    need = ctypes.c_int64()
    nat.call('esr_aggregate_ws_bytes', nat.dtype_code(x.dtype), n, width,
             ctypes.byref(need))
    ws = torch.empty(max(need.value, 1), dtype=torch.uint8, device=x.device)
    out = torch.full((1, width), float('nan'), dtype=x.dtype, device=x.device)
    nat.call('esr_aggregate', op, nat.dtype_code(x.dtype), x.data_ptr(), n, width,
             float(p), out.data_ptr(), ws.data_ptr(), need.value, stream)
    return out


def _noalloc(op, x, p, stream):
    n, width = x.shape
    out = torch.full((1, width), float('nan'), dtype=x.dtype, device=x.device)
    nat.call('esr_aggregate_noalloc', op, nat.dtype_code(x.dtype), x.data_ptr(),
             n, width, float(p), out.data_ptr(), stream)
    return out


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
@pytest.mark.parametrize('width', [1, 3, 16])
def test_noalloc_aggregate_matches_the_workspace_entry_point(gpu, dtype, width):
    g = torch.Generator(device='cpu').manual_seed(width)
    stream = nat.stream_ptr()
    ops = [(nat.AGG_SUM, 2.0), (nat.AGG_NORM, 2.0), (nat.AGG_MAX, 2.0),
           (nat.AGG_NORM, 1.0)]
    for n in (3_000_001, 777):                     # second call: smaller input
        x = (torch.rand((n, width), generator=g, dtype=torch.float64) * 2 - 1) \
            .to(dtype).to(gpu)
        for op, p in ops:
            want = _with_workspace(op, x, p, stream)
            got = _noalloc(op, x, p, stream)
            torch.cuda.synchronize()
            bits = torch.int64 if dtype == torch.float64 else torch.int32
            assert torch.equal(got.view(bits), want.view(bits)), (n, op, p, got, want)
            assert torch.isfinite(got).all()


def test_noalloc_aggregate_twice_on_one_stream_without_sync(gpu):
    """The second call reuses the scratch array while nothing was read back
    in between: ordered on one stream, both results are right."""
    stream = nat.stream_ptr()
    a = torch.arange(1, 2_000_001, dtype=torch.float64, device=gpu).view(-1, 1)
    b = torch.full((5000, 1), 0.5, dtype=torch.float64, device=gpu)
    ra = _noalloc(nat.AGG_SUM, a, 2.0, stream)
    rb = _noalloc(nat.AGG_SUM, b, 2.0, stream)
    torch.cuda.synchronize()
    assert float(ra) == 2_000_000 * 2_000_001 / 2
    assert float(rb) == 2500.0


def test_noalloc_aggregate_rejects_requests_beyond_its_scratch(gpu):
    x = torch.zeros((1 << 20, 64), dtype=torch.float64, device=gpu)
    need = ctypes.c_int64()
    nat.call('esr_aggregate_ws_bytes', nat.F64, x.shape[0], x.shape[1], ctypes.byref(need))
    assert need.value > nat.AGG_SCRATCH_BYTES
    out = torch.zeros((1, 64), dtype=torch.float64, device=gpu)
    with pytest.raises(ValueError, match='scratch'):
        nat.call('esr_aggregate_noalloc', nat.AGG_SUM, nat.F64, x.data_ptr(),
                 x.shape[0], x.shape[1], 2.0, out.data_ptr(), nat.stream_ptr())
