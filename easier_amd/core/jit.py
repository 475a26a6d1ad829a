"""esr.init / esr.compile for the MI355X backend.

Same signatures as easier/core/jit.py:242-419. compile() accepts
backend='hip' (the only backend this package implements) and returns the
same module list with each esr.Module's forward replaced by an engine bound to
GPU-resident, renumbered data.
"""
import os
from typing import List, Optional

import torch

from .. import _native as nat
from . import module as _M
from .compiler import compile_program
from .utils import EasierJitException, init_logger, logger

BACKENDS = ('hip',)


def init(comm_backend: str = 'nccl', **kwargs) -> None:
    """Prepare the runtime. One program drives one GPU; the communication
    backend name is validated for compatibility with easier.init
    (jit.py:242-291) but no process group is needed."""
    if not isinstance(comm_backend, str):
        raise EasierJitException('`comm_backend` must be str')
    world = int(os.environ.get('WORLD_SIZE', '1'))
    if world > 1:
        raise EasierJitException(
            'the hip backend runs one program on one GPU; multi-process '
            f'runs (WORLD_SIZE={world}) are not supported')
    init_logger(0)
    if torch.cuda.is_available():
        torch.cuda.set_device(int(os.environ.get('LOCAL_RANK', '0')))
    logger.info(f"easier_amd initialised (comm_backend={comm_backend!r} "
                "accepted, single GPU)")


def _validate_compile_args(top_modules, backend, partition_mode):
    top_modules = list(top_modules)
    for root in top_modules:
        if not isinstance(root, _M.Module):
            raise EasierJitException(
                f'compile() takes easier.Modules, got {type(root)}')
        if root.easier_jit_backend is not None:
            raise EasierJitException("Input easier.Modules have been compiled.")
        for m in root.modules():
            if isinstance(m, (_M.Selector, _M.Reducer)) and \
                    m.easier_index_status == 'rewritten':
                raise EasierJitException(
                    "Input easier.Modules have been compiled.")
        for p in root.parameters(recurse=True):
            if isinstance(p, _M.Tensor) and p.easier_data_ready:
                raise EasierJitException(
                    "Input easier.Modules have been compiled.")
    if backend not in BACKENDS:
        raise EasierJitException(
            f"Argument `jit_backend` cannot be {backend}; this package "
            "implements backend='hip'")
    if partition_mode not in ('metis', 'evenly'):
        raise EasierJitException(
            f"Argument `partition_mode` cannot be {partition_mode}")
    return top_modules


def compile(modules: List[_M.Module], backend: str = 'hip', *,
            load_dir: Optional[str] = None, partition_mode: str = 'metis',
            reorder: str = 'rcm', fuse: bool = True,
            freeze: bool = False) -> List[_M.Module]:
    """Compile esr.Modules for one MI355X.

    - partition_mode: accepted for compatibility; one GPU holds one partition.
    - reorder: 'rcm' (default) renumbers vertex/cell spaces by reverse
      Cuthill-McKee; 'blocks' into compact runs of 256 (one workgroup tile
      of the row kernels; with freeze=True their gathers then read per-tile
      halos staged in LDS); 'none' keeps the user's order. Results are
      independent of it (Tensor.collect() returns the user's order).
    - fuse: lower supported forwards into fused kernels (hipRTC); otherwise
      every node runs as its own primitive kernel.
    - load_dir: a directory written by esr.dump(); its renumbering is reused
      when it matches this program (core/cache.py), else compiled afresh.
    - freeze: the caller promises that partition tensors no compiled forward
      writes (mesh geometry, coefficients) keep their values after compile();
      fused kernels may then read such a tensor through a compile-time
      dictionary of its distinct values (bit-exact; see core/fuser.py). To
      change one, compile again.
    """
    top_modules = _validate_compile_args(modules, backend, partition_mode)
    cache = None
    if load_dir is not None:
        from .cache import load as load_cache
        cache = load_cache(load_dir)
    if not torch.cuda.is_available():
        raise RuntimeError("backend='hip' needs a visible AMD GPU "
                           "(torch.cuda.is_available() is False)")
    nat.lib()
    device = torch.device('cuda', torch.cuda.current_device())
    logger.info(f"EASIER just-in-time compilation has started, "
                f"backend={backend}, device={device}")
    prog = compile_program(top_modules, reorder, device, cache)
    prog.freeze = bool(freeze)
    from .engine import HipEngine
    fused = None
    if fuse:
        from .fuser import try_fuse_program
        fused = try_fuse_program(prog)
    for m, g in zip(prog.modules, prog.graphs):
        m.easier_jit_backend = backend
        m.partition_mode = partition_mode
        eng = fused.get(m) if fused else None
        if eng is None:
            eng = HipEngine(m, g, device)
        m.easier_engine = eng
        m.easier_program, m.easier_graph = prog, g
        m.forward = eng.forward
    for m in prog.modules:
        hook = getattr(m, 'easier_after_compile', None)
        if hook is not None:
            hook(prog)
    logger.info("EASIER just-in-time compilation has completed")
    return top_modules
