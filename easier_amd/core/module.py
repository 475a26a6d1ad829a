"""The EASIER data model: esr.Tensor, esr.Selector, esr.Reducer, esr.Module,
the aggregators and the tensor factories.

Same names, constructor arguments, error types and messages as
easier/core/module.py (Selector :339-383, Reducer :389-477, Tensor :480-707,
Module :710-822, aggregators :825-884, factories :35-243). After
esr.compile(..., backend='hip') every primitive executes on the GPU through
libesrhip.so; before it, Selector/Reducer raise like the reference.
"""
import builtins
from typing import List, Optional, Sequence, Tuple, Union

import torch
from torch import nn
from typing_extensions import Literal, Self

from . import ops
from .dataloader import (ATTRIBUTE_PLACEHOLDER, ArangeTensorLoader,
                         DataLoaderBase, FulledTensorLoader, H5DataLoader,
                         InMemoryTensorLoader)
from .utils import EasierJitException


# ---------------------------------------------------------------------------
# factories (module.py:35-243)
# ---------------------------------------------------------------------------
def hdf5(file: str, dataset: str, dtype: Optional[torch.dtype] = None,
         device: Union[torch.device, str, None] = None, **h5_file_kwargs):
    return H5DataLoader(file, dataset, dtype=dtype,
                        device=device if device is not None else 'cpu',
                        **h5_file_kwargs)


def full(size: Sequence[int], fill_value, *, dtype=None, device=None):
    if isinstance(fill_value, bool) or not isinstance(fill_value, (int, float)):
        raise TypeError('fill_value must be integer or floating-point')
    if dtype is None:
        dtype = torch.int64 if isinstance(fill_value, int) else torch.float64
    return FulledTensorLoader(fill_value, size, dtype,
                              device if device is not None else 'cpu')


def zeros(size: Sequence[int], dtype=None, device=None):
    return full(size, 0, dtype=dtype or torch.float64, device=device)


def ones(size: Sequence[int], dtype=None, device=None):
    return full(size, 1, dtype=dtype or torch.float64, device=device)


def _dtype_device_like(input, dtype, device):
    return (dtype if dtype is not None else input.dtype,
            torch.device(device if device is not None else input.device))


def full_like(input, fill_value, *, dtype=None, device=None):
    dtype, device = _dtype_device_like(input, dtype, device)
    return full(input.shape, fill_value, dtype=dtype, device=device)


def zeros_like(input, dtype=None, device=None):
    dtype, device = _dtype_device_like(input, dtype, device)
    return zeros(input.shape, dtype=dtype, device=device)


def ones_like(input, dtype=None, device=None):
    dtype, device = _dtype_device_like(input, dtype, device)
    return ones(input.shape, dtype=dtype, device=device)


def arange(*args, **kwargs):
    def _end(end, dtype=None, device=None):
        return (0, end, 1, dtype, device)

    def _start_end(start, end, step=1, dtype=None, device=None):
        return (start, end, step, dtype, device)

    for matcher in (_end, _start_end):
        try:
            start, end, step, dtype, device = matcher(*args, **kwargs)
            break
        except TypeError:
            continue
    else:
        raise TypeError(f'Unexpected arguments {args} to easier.arange')
    for a in (start, end, step):
        if not isinstance(a, (int, float)):
            raise TypeError(
                'argument to easier.arange must be integer or floating-point')
    if dtype is None:
        dtype = torch.int64 if isinstance(start + end + step, int) \
            else torch.float64
    return ArangeTensorLoader(start, end, step, dtype,
                              device if device is not None else 'cpu')


def _resolve_data_loader(arg) -> DataLoaderBase:
    if isinstance(arg, DataLoaderBase):
        return arg
    if isinstance(arg, Tensor) or hasattr(arg, ATTRIBUTE_PLACEHOLDER):
        raise TypeError(
            "Cannot use another easier.Tensor (or its .data property)"
            " to initialize easier.Tensor")
    if isinstance(arg, torch.Tensor):
        return InMemoryTensorLoader(arg)
    raise TypeError(f"Unknown data type {type(arg)}")


def _resolve_to_args(cls_name, args, kwargs):
    vals = set(args).union(kwargs.values())
    if len(vals) != 1:
        raise NotImplementedError(
            f'Currently easier.{cls_name}.to() only supports one argument,'
            ' consider using method chaining `.to(dtype).to(device)`')
    (arg,) = vals
    if isinstance(arg, torch.dtype):
        return None, arg
    if isinstance(arg, (str, torch.device)):
        return torch.device(arg), None
    raise TypeError(f"easier.{cls_name}.to() does not accept the argument"
                    f" of type {type(arg)}")


# ---------------------------------------------------------------------------
# Selector / Reducer
# ---------------------------------------------------------------------------
class Selector(nn.Module):
    """out = tensor[idx] (module.py:339-383)."""

    def __init__(self, idx: Union[torch.Tensor, DataLoaderBase]):
        super().__init__()
        self.easier_data_loader = _resolve_data_loader(idx)
        self.easier_index_status = 'placeholder'
        self.idx: torch.Tensor = self.easier_data_loader.get_placeholder()
        self.easier_hint_name: str = ''
        # Set by compile(): the device index (int32, renumbered) and the
        # exclusive bound every input's leading dimension must exceed.
        self.easier_idx_max: int = -1

    def forward(self, tensor: torch.Tensor) -> torch.Tensor:
        if self.easier_index_status != 'rewritten':
            raise RuntimeError("Selector.idx not ready, run compile() first")
        if tensor.shape[0] <= self.easier_idx_max:
            raise IndexError(
                f'{self.easier_hint_name}: index {self.easier_idx_max} is out '
                f'of bounds for a tensor of length {tensor.shape[0]}')
        return ops.gather(tensor, self.idx)

    def to(self, device: Union[str, torch.device]) -> Self:
        if self.easier_index_status != 'placeholder':
            raise RuntimeError(
                "The properties of easier.Selector can only be modified"
                " before easier.compile()")
        self.easier_data_loader = self.easier_data_loader.to(device=device)
        self.idx = self.easier_data_loader.get_placeholder()
        return self


ReduceOp = Literal['sum', 'prod', 'mean', 'amax', 'amin']


class Reducer(nn.Module):
    """out = zeros(n, ...).scatter_reduce_(0, idx, tensor, reduce,
    include_self=False) (module.py:389-477), evaluated as a CSR segmented
    reduction in edge order."""

    def __init__(self, idx: Union[torch.Tensor, DataLoaderBase], n: int,
                 reduce: ReduceOp = 'sum') -> None:
        super().__init__()
        if reduce not in ('sum', 'prod', 'mean', 'amax', 'amin'):
            raise ValueError(f'unknown reduce op {reduce!r}')
        self.n = n
        self.reduce = reduce
        self.easier_data_loader = _resolve_data_loader(idx)
        self.easier_index_status = 'placeholder'
        self.idx: torch.Tensor = self.easier_data_loader.get_placeholder()
        self.easier_hint_name: str = ''
        # Set by compile(): CSR over the (renumbered) targets.
        self.easier_rowptr: Optional[torch.Tensor] = None
        self.easier_perm: Optional[torch.Tensor] = None
        self.easier_long_rows: Optional[torch.Tensor] = None

    def forward(self, tensor: torch.Tensor, *,
                out: Optional[torch.Tensor] = None) -> torch.Tensor:
        if self.easier_index_status != 'rewritten':
            raise RuntimeError("Reducer.idx not ready, run compile() first")
        if tensor.shape[0] != self.idx.shape[0]:
            raise ValueError(
                f'{self.easier_hint_name}: input length {tensor.shape[0]} does'
                f' not match the index length {self.idx.shape[0]}')
        if out is not None:
            if out.shape[0] != self.n:
                raise ValueError(
                    f"{out.shape[0]} the length of the first dimension of "
                    f"`out` and {self.n} the specified `Reducer.n` do not "
                    "match")
            if tuple(out.shape[1:]) != tuple(tensor.shape[1:]):
                raise ValueError('`out` must have the shape (n,) + '
                                 'tensor.shape[1:]')
        return ops.csr_reduce(tensor, self.easier_rowptr, self.easier_perm,
                              self.easier_long_rows, self.n, self.reduce,
                              out=out, keep_empty=out is not None)

    def to(self, device: Union[str, torch.device]) -> Self:
        if self.easier_index_status != 'placeholder':
            raise RuntimeError(
                "The properties of easier.Reducer can only be modified"
                " before easier.compile()")
        self.easier_data_loader = self.easier_data_loader.to(device=device)
        self.idx = self.easier_data_loader.get_placeholder()
        return self


# ---------------------------------------------------------------------------
# Tensor
# ---------------------------------------------------------------------------
class Tensor(nn.Parameter):
    """A mesh-distributed ('partition') or replicated ('replicate') tensor
    (module.py:480-707)."""

    def __new__(cls, data, mode: Literal['partition', 'replicate'],
                requires_grad: bool = False) -> "Tensor":
        dl = _resolve_data_loader(data)
        t = super().__new__(cls, dl.get_placeholder(), requires_grad)
        t.easier_data_loader = dl
        return t

    def __init__(self, data, mode: Literal['partition', 'replicate'],
                 requires_grad: bool = False) -> None:
        if mode not in ('partition', 'replicate'):
            raise ValueError(f'unknown mode {mode!r}')
        self.easier_data_ready = False
        self.is_partition = mode == 'partition'
        self.is_replica = mode == 'replicate'
        self.easier_hint_name = ''
        # Set by compile() for partition tensors: order[new] = old position.
        self.easier_order: Optional[torch.Tensor] = None

    def __repr__(self) -> str:
        if self.is_replica:
            body = (' ' + repr(self.easier_data_loader)) \
                if not self.easier_data_ready else '\n' + repr(self.data)
            return 'Replicated easier.Tensor' + body
        body = '' if self.easier_data_ready else \
            ' ' + repr(self.easier_data_loader)
        return 'Managed partitioned easier.Tensor' + body

    def __setitem__(self, indices, val) -> Self:
        if not self.easier_data_ready:
            raise RuntimeError("Tensor data is not ready, run compile() first")
        ops.copy_(self.data[indices], val)
        return self

    def to(self, *args, **kwargs) -> Self:
        if self.easier_data_ready is not False:
            raise RuntimeError(
                "The properties of easier.Tensor can only be modified"
                " before easier.compile()")
        device, dtype = _resolve_to_args('Tensor', args, kwargs)
        self.easier_data_loader = self.easier_data_loader.to(
            device=device, dtype=dtype)
        self.data = self.easier_data_loader.get_placeholder()
        return self

    def collect(self) -> torch.Tensor:
        """The full tensor in the user's original element order, on the
        device the program declared (module.py:588-612)."""
        if not self.easier_data_ready:
            raise RuntimeError("Tensor data is not ready, run compile() first")
        host = self.data.detach().cpu()
        if self.easier_order is not None:
            full_t = torch.empty_like(host)
            full_t[self.easier_order] = host
            host = full_t
        return host.to(self.easier_data_loader.device, copy=True)

    def save(self, h5_file_path, h5_dataset_path, **h5_file_kwargs) -> None:
        """Add the tensor (in the user's order) to an HDF5 file as a new
        dataset, creating the file if needed (module.py:614-707, h5py mode
        'a'), with the package's own HDF5 writer (core/h5.py); an existing
        dataset of that name raises. h5_file_kwargs are accepted and ignored."""
        if not self.easier_data_ready:
            raise RuntimeError("Tensor data is not ready, run compile() first")
        import os
        from . import h5
        path = os.path.expanduser(h5_file_path)
        d = os.path.dirname(path)
        if d:
            os.makedirs(d, exist_ok=True)
        h5.append(path, {h5_dataset_path: self.collect().cpu().numpy()})


# ---------------------------------------------------------------------------
# Module
# ---------------------------------------------------------------------------
class Module(nn.Module):
    """A mesh program: forward() takes no arguments and returns None
    (module.py:710-822)."""

    def __init__(self):
        super().__init__()
        self.easier_jit_backend: Optional[str] = None
        self.easier_hint_name: str = ''
        self.partition_mode = 'metis'

    def forward(self):
        pass

    def to(self, *args, **kwargs) -> Self:
        if self.easier_jit_backend is not None:
            raise RuntimeError(
                "The properties of easier.Module can only be modified"
                " before easier.compile()")
        device, fp_dtype = _resolve_to_args('Module', args, kwargs)
        if fp_dtype is not None and not fp_dtype.is_floating_point:
            raise ValueError(
                f"easier.Module.to() does not accept the dtype {fp_dtype},"
                " floating point dtypes are required")
        from .compiler import get_easier_objects
        for obj in get_easier_objects([self]):
            if isinstance(obj, (Selector, Reducer)) and device is not None:
                obj.to(device)
            if isinstance(obj, Tensor):
                if device is not None:
                    obj.to(device)
                if fp_dtype is not None and \
                        obj.easier_data_loader.dtype.is_floating_point:
                    obj.to(fp_dtype)
            if isinstance(obj, Module):
                for name, dl in list(obj.__dict__.items()):
                    if isinstance(dl, DataLoaderBase):
                        dt = fp_dtype if dl.dtype.is_floating_point else None
                        setattr(obj, name, dl.to(dtype=dt, device=device))
        return self


# ---------------------------------------------------------------------------
# aggregators (module.py:825-884): reduce dim 0 with keepdim=True
# ---------------------------------------------------------------------------
def sum(tensor: torch.Tensor) -> torch.Tensor:
    """torch.sum(tensor, dim=0, keepdim=True)."""
    return ops.aggregate('sum', tensor)


def prod(tensor: torch.Tensor) -> torch.Tensor:
    """torch.prod(tensor, dim=0, keepdim=True)."""
    return ops.aggregate('prod', tensor)


def norm(tensor: torch.Tensor, p: Union[int, str] = 2) -> torch.Tensor:
    """torch.norm(tensor, p, dim=0, keepdim=True)."""
    return ops.aggregate('norm', tensor, p=p)


def max(tensor: torch.Tensor) -> torch.Tensor:
    """torch.amax(tensor, dim=0, keepdim=True)."""
    return ops.aggregate('max', tensor)


def min(tensor: torch.Tensor) -> torch.Tensor:
    """torch.amin(tensor, dim=0, keepdim=True)."""
    return ops.aggregate('min', tensor)


easier_aggregators = (sum, prod, norm, max, min)
