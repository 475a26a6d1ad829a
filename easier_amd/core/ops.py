"""Device operations on torch tensors, all executed by libesrhip.so.

torch is used here only for what it does without computing: allocation
(`torch.empty`), views (`expand`, slicing), dtype promotion rules
(`torch.result_type`) and shape broadcasting (`torch.broadcast_shapes`).
"""
import ctypes
import numbers
from typing import Optional

import torch

from .. import _native as nat

_MAXDIM = 6


def _dev_check(*ts):
    for t in ts:
        if isinstance(t, torch.Tensor) and t.device.type != 'cuda':
            raise RuntimeError(
                f'hip backend operation got a tensor on {t.device}; '
                'compiled data lives on the GPU')


# ---------------------------------------------------------------------------
# allocation helpers
# ---------------------------------------------------------------------------
def empty(shape, dtype, device):
    return torch.empty(tuple(shape), dtype=dtype, device=device)


def fill_(t: torch.Tensor, value) -> torch.Tensor:
    """t[...] = value (any layout)."""
    _dev_check(t)
    if t.numel() == 0:
        return t
    if t.is_contiguous():
        nat.call('esr_fill', nat.dtype_code(t.dtype), t.data_ptr(), t.numel(),
                 float(value), nat.stream_ptr())
        return t
    return ewise('copy', value, out=t)


def zeros(shape, dtype, device):
    return fill_(empty(shape, dtype, device), 0)


# ---------------------------------------------------------------------------
# elementwise
# ---------------------------------------------------------------------------
def _is_scalar(x):
    return isinstance(x, (numbers.Number, bool)) and not isinstance(x, torch.Tensor)


def _float_default(dtype):
    return dtype if dtype.is_floating_point else torch.get_default_dtype()


def _compute_dtype(op, operands):
    vals = list(operands[1:]) if op == 'where' else list(operands)
    if len(vals) == 1:
        (x,) = vals
        if not isinstance(x, torch.Tensor):
            raise RuntimeError(f'elementwise {op} of a Python scalar')
        dt = x.dtype
    else:
        dt = torch.result_type(vals[0], vals[1])
    if op in ('div', 'exp', 'log', 'sqrt', 'rsqrt', 'sin', 'cos', 'tanh',
              'sigmoid', 'reciprocal', 'atan2'):
        dt = _float_default(dt)
    if op == 'pow' and not dt.is_floating_point:
        dt = _float_default(dt) if any(isinstance(v, float) for v in vals) else dt
    return dt


def _cast_to(t: torch.Tensor, dtype):
    if t.dtype == dtype:
        return t
    out = empty(t.shape, dtype, t.device)
    _launch('copy', [t], out, t.dtype)
    return out


def _launch(op, inputs, out, in_dtype, scalars=None):
    shape = tuple(out.shape)
    ndim = len(shape)
    if ndim > _MAXDIM:
        raise RuntimeError(f'elementwise operand of rank {ndim} > {_MAXDIM}')
    n_in = len(inputs)
    ptrs = (ctypes.c_void_p * 3)()
    istr = (ctypes.c_int64 * (3 * max(ndim, 1)))()
    sc = (ctypes.c_double * 3)()
    for k, x in enumerate(inputs):
        if isinstance(x, torch.Tensor):
            xe = x.expand(shape) if tuple(x.shape) != shape else x
            ptrs[k] = xe.data_ptr()
            for d, s in enumerate(xe.stride()):
                istr[k * ndim + d] = s
        else:
            ptrs[k] = None
            sc[k] = float(x)
    if scalars:
        for k, v in scalars.items():
            sc[k] = float(v)
    shp = (ctypes.c_int64 * max(ndim, 1))(*shape)
    ostr = (ctypes.c_int64 * max(ndim, 1))(*out.stride())
    nat.call('esr_ewise', nat.EW[op], ndim, shp, nat.dtype_code(out.dtype),
             out.data_ptr(), ostr, nat.dtype_code(in_dtype), n_in, ptrs, istr,
             sc, nat.stream_ptr())


def ewise(op: str, *operands, out: Optional[torch.Tensor] = None):
    """Evaluate one elementwise torch operator with torch's broadcasting and
    type promotion. `out` (if given) must have the broadcast shape."""
    tensors = [x for x in operands if isinstance(x, torch.Tensor)]
    _dev_check(*tensors, out)
    if not tensors and out is None:
        raise RuntimeError(f'elementwise {op} without a tensor operand')
    device = (tensors[0] if tensors else out).device
    if op == 'copy':
        (src,) = operands
        if out is None:
            raise RuntimeError('copy needs a destination')
        in_dtype = src.dtype if isinstance(src, torch.Tensor) else out.dtype
        _launch('copy', [src], out, in_dtype)
        return out
    cdt = _compute_dtype(op, operands)
    if op == 'pow' and not cdt.is_floating_point and _is_scalar(operands[1]) and \
            operands[1] < 0:
        raise RuntimeError('Integers to negative integer powers are not allowed.')
    shape = torch.broadcast_shapes(*[tuple(t.shape) for t in tensors]) \
        if tensors else tuple(out.shape)
    inputs = []
    for k, x in enumerate(operands):
        if isinstance(x, torch.Tensor):
            want = torch.bool if (op == 'where' and k == 0) else cdt
            if op == 'where' and k == 0 and x.dtype != torch.bool:
                x = ewise('ne', x, 0)
            elif x.dtype != want:
                x = _cast_to(x, want)
            inputs.append(x)
        else:
            inputs.append(x)
    if op in nat.EW_PREDICATES:
        odt = torch.bool
    else:
        odt = cdt
    if out is None:
        out = empty(shape, odt, device)
    else:
        if tuple(out.shape) != tuple(shape):
            shape_b = torch.broadcast_shapes(tuple(out.shape), shape)
            if shape_b != tuple(out.shape):
                raise RuntimeError(
                    f'output of shape {tuple(out.shape)} cannot hold the '
                    f'broadcast shape {shape}')
        if out.dtype != odt:
            tmp = ewise(op, *operands)
            return ewise('copy', tmp, out=out)
    _launch(op, inputs, out, cdt)
    return out


def copy_(dst: torch.Tensor, src) -> torch.Tensor:
    """dst[...] = src, broadcasting src and casting to dst.dtype."""
    _dev_check(dst, src if isinstance(src, torch.Tensor) else None)
    if isinstance(src, torch.Tensor):
        shape_b = torch.broadcast_shapes(tuple(dst.shape), tuple(src.shape))
        if shape_b != tuple(dst.shape):
            raise RuntimeError(
                f'cannot assign shape {tuple(src.shape)} into {tuple(dst.shape)}')
        _launch('copy', [src], dst, src.dtype)
    else:
        _launch('copy', [src], dst, dst.dtype)
    return dst


def contiguous(t: torch.Tensor) -> torch.Tensor:
    if t.is_contiguous():
        return t
    out = empty(t.shape, t.dtype, t.device)
    return copy_(out, t)


# ---------------------------------------------------------------------------
# Selector / Reducer / aggregators
# ---------------------------------------------------------------------------
def gather(src: torch.Tensor, idx: torch.Tensor, out=None) -> torch.Tensor:
    """out[e] = src[idx[e]] (rows of src.shape[1:])."""
    _dev_check(src, idx)
    src = contiguous(src)
    n_out = idx.shape[0]
    width = 1
    for s in src.shape[1:]:
        width *= s
    if out is None:
        out = empty((n_out,) + tuple(src.shape[1:]), src.dtype, src.device)
    if width == 0 or n_out == 0:
        return out
    nat.call('esr_gather', nat.dtype_code(src.dtype), src.data_ptr(),
             src.shape[0], width, idx.data_ptr(), n_out, out.data_ptr(),
             nat.stream_ptr())
    return out


def csr_reduce(vals: torch.Tensor, rowptr, perm, long_rows, n_rows, op,
               out=None, keep_empty=False) -> torch.Tensor:
    _dev_check(vals)
    vals = contiguous(vals)
    width = 1
    for s in vals.shape[1:]:
        width *= s
    if out is None:
        out = empty((n_rows,) + tuple(vals.shape[1:]), vals.dtype, vals.device)
        keep_empty = False
    if width == 0 or n_rows == 0:
        return out
    n_long = 0 if long_rows is None else long_rows.shape[0]
    target = out if out.is_contiguous() else empty(out.shape, out.dtype, out.device)
    if target is not out and keep_empty:
        copy_(target, out)
    nat.call('esr_csr_reduce', nat.RED_OPS[op], nat.dtype_code(vals.dtype),
             vals.data_ptr(), width, rowptr.data_ptr(),
             None if perm is None else perm.data_ptr(), n_rows,
             None if n_long == 0 else long_rows.data_ptr(), n_long,
             1 if keep_empty else 0, target.data_ptr(), nat.stream_ptr())
    if target is not out:
        copy_(out, target)
    return out


_AGG = {'sum': nat.AGG_SUM, 'prod': nat.AGG_PROD, 'norm': nat.AGG_NORM,
        'max': nat.AGG_MAX, 'min': nat.AGG_MIN}


def aggregate_ws_bytes(t: torch.Tensor) -> int:
    """Bytes of the workspace esr_aggregate needs for `t`."""
    width = 1
    for s in t.shape[1:]:
        width *= s
    nb = ctypes.c_int64(0)
    nat.call('esr_aggregate_ws_bytes', nat.dtype_code(t.dtype),
             t.shape[0] if t.dim() > 0 else 1, width, ctypes.byref(nb))
    return max(nb.value, 8)


def aggregate(op: str, t: torch.Tensor, p=2.0, keep_leading=True, out=None, ws=None):
    """torch.<op>(t, dim=0, keepdim=True) (norm: p-norm) as a deterministic
    two-stage device reduction. With `out` and `ws` (aggregate_ws_bytes
    bytes) given, nothing is allocated (safe inside a captured graph)."""
    _dev_check(t)
    if op == 'norm':
        if isinstance(p, str):
            p = {'fro': 2.0, 'inf': float('inf'), '-inf': float('-inf')}[p]
        p = float(p)
        if not t.dtype.is_floating_point:
            raise RuntimeError('norm needs a floating-point tensor')
    t = contiguous(t)
    n = t.shape[0] if t.dim() > 0 else 1
    width = 1
    for s in t.shape[1:]:
        width *= s
    out_shape = ((1,) + tuple(t.shape[1:])) if keep_leading else tuple(t.shape[1:])
    if out is None:
        out = empty(out_shape, t.dtype, t.device)
    elif out.numel() != max(width, 1) or out.dtype != t.dtype or not out.is_contiguous():
        raise ValueError('aggregate: `out` must be a contiguous tensor of the result size')
    if width == 0:
        return out
    if ws is None:
        ws = torch.empty((aggregate_ws_bytes(t),), dtype=torch.uint8, device=t.device)
    elif ws.numel() * ws.element_size() < aggregate_ws_bytes(t):
        raise ValueError('aggregate: workspace too small')
    nat.call('esr_aggregate', _AGG[op], nat.dtype_code(t.dtype), t.data_ptr(),
             n, width, float(p), out.data_ptr(), ws.data_ptr(), ws.numel(),
             nat.stream_ptr())
    return out


def full_reduce(op: str, t: torch.Tensor, p=2.0):
    """torch.<op>(t) over every element -> 0-dim tensor."""
    flat = contiguous(t).reshape(-1, 1) if t.dim() > 0 else t.reshape(1, 1)
    out = aggregate(op, flat, p=p, keep_leading=True)
    return out.reshape(())


def gemv_rows(A: torch.Tensor, x: torch.Tensor, out=None):
    """out[r] = sum_k A[r, k] * x[k] (A (n, k) any strides, x (k,) or (k, 1))."""
    _dev_check(A, x)
    n, kdim = A.shape
    dt = torch.result_type(A, x)
    A = _cast_to(A, dt)
    x = _cast_to(x, dt)
    xs = x.stride(0) if x.dim() >= 1 and x.shape[0] > 0 else 1
    if out is None:
        out = empty((n,), dt, A.device)
    nat.call('esr_gemv_rows', nat.dtype_code(dt), A.data_ptr(), n, kdim,
             A.stride(0), A.stride(1), x.data_ptr(), xs, out.data_ptr(),
             nat.stream_ptr())
    return out
