"""Selector, Reducer and aggregators as the reference evaluates them on CPU.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
"""
import torch


def selector(t: torch.Tensor, idx: torch.Tensor) -> torch.Tensor:
    """easier/core/module.py:364-371: tensor[idx]."""
    return t[idx]


def reducer(t: torch.Tensor, idx: torch.Tensor, n: int, reduce='sum',
            out=None) -> torch.Tensor:
    """easier/core/module.py:441-461: zeros((n,)+shape[1:]) (or `out`), then
    scatter_reduce_(0, idx expanded, t, reduce, include_self=False)."""
    shape = t.shape
    if out is None:
        out = torch.zeros((n,) + tuple(shape[1:]), dtype=t.dtype)
    if out.shape[0] != n:
        raise ValueError('out length mismatch')
    idx = idx.to(torch.int64)
    ie = idx[(...,) + (None,) * (len(shape) - 1)].expand(-1, *shape[1:])
    return out.scatter_reduce_(0, ie, t, reduce, include_self=False)


def reducer_loop(t: torch.Tensor, idx, n: int, reduce='sum', out=None):
    """Brute-force restatement of the same reduction: each target row starts
    from the op's identity and folds its entries in index order; rows without
    entries keep `out` (zeros without it). Pure Python: small inputs only."""
    import math
    rows = {}
    for e, r in enumerate(idx.tolist()):
        rows.setdefault(r, []).append(e)
    res = torch.zeros((n,) + tuple(t.shape[1:]), dtype=t.dtype) if out is None \
        else out.clone()
    flat = t.reshape(t.shape[0], -1)
    resf = res.reshape(n, -1)
    for r, es in rows.items():
        for c in range(flat.shape[1]):
            if reduce in ('sum', 'mean'):
                acc = torch.zeros((), dtype=t.dtype)
            elif reduce == 'prod':
                acc = torch.ones((), dtype=t.dtype)
            elif reduce == 'amax':
                acc = torch.tensor(-math.inf, dtype=t.dtype)
            else:
                acc = torch.tensor(math.inf, dtype=t.dtype)
            for e in es:
                v = flat[e, c]
                if reduce in ('sum', 'mean'):
                    acc = acc + v
                elif reduce == 'prod':
                    acc = acc * v
                elif reduce == 'amax':
                    acc = v if (torch.isnan(v) or v > acc) else acc
                else:
                    acc = v if (torch.isnan(v) or v < acc) else acc
            if reduce == 'mean':
                acc = acc / len(es)
            resf[r, c] = acc
    return res


def aggregate(op: str, t: torch.Tensor, p=2) -> torch.Tensor:
    """easier/core/module.py:825-884 at one partition: torch.<op>(t, dim=0,
    keepdim=True); the all-gather + second reduction of
    dataflow_distribution.py:197-289 is an identity on one (1, ...) row."""
    if op == 'sum':
        return torch.sum(t, dim=0, keepdim=True)
    if op == 'prod':
        return torch.prod(t, dim=0, keepdim=True)
    if op == 'norm':
        return torch.norm(t, p, dim=0, keepdim=True)
    if op == 'max':
        return torch.amax(t, dim=0, keepdim=True)
    if op == 'min':
        return torch.amin(t, dim=0, keepdim=True)
    raise ValueError(op)
