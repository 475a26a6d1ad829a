"""Data sources behind esr.Tensor / esr.Selector / esr.Reducer.

Follows the roles of easier/core/runtime/data_loader.py (DataLoaderBase
:61-313, InMemoryTensorLoader :316-403, FulledTensorLoader :754-830,
ArangeTensorLoader :833-931): a loader is a lightweight description
(shape, dtype, device) until esr.compile() loads it. There is a single
partition, so a loader only ever loads fully, onto the host, and the compiler
permutes and uploads the result.
"""
import copy
import os
from typing import Optional, Tuple, Union

import numpy as np
import torch

from .utils import EasierJitException

ATTRIBUTE_PLACEHOLDER = "easier_placeholder"


class DataLoaderBase:
    shape: Tuple[int, ...]
    dtype: torch.dtype
    device: torch.device

    def to(self, *, dtype: Optional[torch.dtype] = None,
           device: Optional[Union[torch.device, str]] = None
           ) -> 'DataLoaderBase':
        clone = copy.copy(self)
        if dtype is not None:
            clone.dtype = dtype
        if device is not None:
            clone.device = torch.device(device)
        return clone

    def load_cpu(self) -> torch.Tensor:
        """The full tensor on the host, in self.dtype."""
        raise NotImplementedError

    def minmax(self):
        t = self.load_cpu()
        if t.numel() == 0:
            return 0, -1
        return t.min().item(), t.max().item()

    def get_placeholder(self, device=None) -> torch.Tensor:
        """A zero-memory stand-in with the right shape/dtype
        (data_loader.py:249-300): a size-1 buffer expanded to `shape`."""
        if device is None:
            device = self.device
        if len(self.shape) == 0:
            ph = torch.zeros((), dtype=self.dtype, device=device)
        else:
            ph = torch.zeros((1,), dtype=self.dtype,
                             device=device).expand(self.shape)

        def _to_forbidden(self, *args, **kwargs):
            raise EasierJitException("Cannot call .to() on placeholders!")
        ph.to = _to_forbidden.__get__(ph)
        setattr(ph, ATTRIBUTE_PLACEHOLDER, True)
        return ph


class InMemoryTensorLoader(DataLoaderBase):
    def __init__(self, tensor: torch.Tensor) -> None:
        self.dtype = tensor.dtype
        self.shape = tuple(tensor.shape)
        self.device = tensor.device
        self.tensor = tensor.detach().cpu()

    def load_cpu(self):
        return self.tensor.to(self.dtype)

    def __repr__(self):
        return f'{self.__class__.__name__}(tensor={self.tensor})'


class FulledTensorLoader(DataLoaderBase):
    def __init__(self, value, shape, dtype, device) -> None:
        self.value = value
        self.shape = tuple(shape)
        self.dtype = dtype
        self.device = torch.device(device)

    def load_cpu(self):
        return torch.full(self.shape, self.value, dtype=self.dtype)

    def minmax(self):
        return self.value, self.value

    def __repr__(self):
        return (f'{self.__class__.__name__}(value={self.value}, '
                f'shape={self.shape}, dtype={self.dtype})')


class ArangeTensorLoader(DataLoaderBase):
    def __init__(self, start, end, step, dtype, device) -> None:
        if step == 0:
            raise ValueError("step must not be 0")
        self.start, self.end, self.step = start, end, step
        self.dtype = dtype
        self.device = torch.device(device)
        length = len(range(0, int(np.ceil((end - start) / step))))
        self.shape = (max(0, length),)

    def load_cpu(self):
        return torch.arange(self.start, self.end, self.step, dtype=self.dtype)

    def __repr__(self):
        return (f'{self.__class__.__name__}(start={self.start}, '
                f'end={self.end}, step={self.step}, dtype={self.dtype})')


class H5DataLoader(DataLoaderBase):
    """A dataset in an HDF5 file (data_loader.py:418-751; e.g. the mesh schema
    of tutorial/create_triangular_mesh.py:29-35), read with the package's own
    HDF5 reader (core/h5.py: contiguous or compact datasets of integers or
    floating point, as h5py writes them by default) -- h5py is not needed.
    Shape and dtype are read when the loader is created; the values when the
    program is compiled. h5_file_kwargs (h5py.File options in the reference)
    are accepted and ignored."""

    def __init__(self, file: str, dataset: str, dtype=None, device='cpu',
                 **h5_file_kwargs) -> None:
        from . import h5
        self.file, self.dataset = os.path.expanduser(file), dataset
        self.h5_file_kwargs = h5_file_kwargs
        with h5.H5Reader(self.file) as r:
            d = r.dataset(dataset)
        self.shape = tuple(d.shape)
        raw = torch.from_numpy(np.empty((0,), dtype=d.dtype.newbyteorder('='))).dtype
        self.dtype = dtype if dtype is not None else raw
        self.device = torch.device(device)

    def load_rows(self, start: int, stop: int) -> torch.Tensor:
        """Rows [start, stop) only (partially_load_by_chunk's unit, :572-590)."""
        from . import h5
        with h5.H5Reader(self.file) as r:
            arr = r.read(self.dataset, start, stop)
        return torch.from_numpy(np.ascontiguousarray(arr)).to(self.dtype)

    def load_cpu(self):
        from . import h5
        with h5.H5Reader(self.file) as r:
            arr = r.read(self.dataset)
        return torch.from_numpy(np.ascontiguousarray(arr)).to(self.dtype)

    def __repr__(self):
        return (f'{self.__class__.__name__}(file={self.file!r}, '
                f'dataset={self.dataset!r}, dtype={self.dtype})')
