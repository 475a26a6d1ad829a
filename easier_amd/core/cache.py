"""Compile cache: esr.dump() and esr.compile(..., load_dir=...).

Mirrors the reference's contract (easier/core/dump.py:341-503 `dump`,
:824-918 `load_dumps`, jit.py:385-395): dump() writes the compilation state
of compiled modules under `dump_dir/jit/`; compile(load_dir=dump_dir) reuses
it when it matches the program being compiled and otherwise compiles from
scratch (the reference likewise falls back when the global configuration
changed and invalidates the dump when the traced program changed,
dump.py:921-999).

What is cached is what the renumbering computes: the order of every
renumbered index space (reverse Cuthill-McKee for vertex/cell spaces, the
stable Reducer sort for edge spaces). A space is named by the smallest of
its members' names ('<hint name>:t' for a tensor, ':in'/':out' for a
Selector's or Reducer's input and output), which depends only on the
program's structure. The dump is valid for a program whose fingerprint --
traced graph of every module, every Selector/Reducer's name, size and index
contents, every tensor's name and shape, the reorder mode and the ABI
version -- equals the dumped one.

Files: `jit/hip_0.json` (fingerprint, space names and kinds: 'rcm' for a
renumbered vertex/cell space, 'edge' for a sorted edge space) and
`jit/hip_0.npz` (one int64 order per space), read with json and numpy.load(allow_pickle=False):
loading executes nothing from the files. (The reference writes
`jit/jit_<rank>.hdf5`; h5py is not a dependency here.)
"""
import hashlib
import json
import os
from typing import Dict, List, Optional

import numpy as np

from . import module as _M
from .utils import EasierJitException, logger

FORMAT = 1
JSON_NAME = 'hip_0.json'
NPZ_NAME = 'hip_0.npz'


def _sha(a: np.ndarray) -> str:
    a = np.ascontiguousarray(a)
    return hashlib.sha256(str(a.dtype).encode() + str(a.shape).encode() +
                          a.tobytes()).hexdigest()


def space_names(sa, objs) -> Dict[int, str]:
    """Space root -> the smallest member name of the space."""
    members: Dict[int, List[str]] = {}

    def add(var, name):
        members.setdefault(sa.root(var), []).append(name)
    for o in objs:
        name = o.easier_hint_name
        if isinstance(o, _M.Tensor) and o in sa.tensor_var:
            add(sa.tensor_var[o], name + ':t')
        elif isinstance(o, _M.Selector) and o in sa.sel_vars:
            add(sa.sel_vars[o][0], name + ':in')
            add(sa.sel_vars[o][1], name + ':out')
        elif isinstance(o, _M.Reducer) and o in sa.red_vars:
            add(sa.red_vars[o][0], name + ':in')
            add(sa.red_vars[o][1], name + ':out')
    return {root: min(names) for root, names in members.items()}


def fingerprint(modules, graphs, objs, sel_idx, red_idx, reorder, abi) -> dict:
    fp = {'format': FORMAT, 'abi': int(abi), 'reorder': reorder, 'world_size': 1,
          'modules': [], 'selectors': {}, 'reducers': {}, 'tensors': {}}
    for m, g in zip(modules, graphs):
        code = g.python_code('self').src
        fp['modules'].append([m.easier_hint_name,
                              hashlib.sha256(code.encode()).hexdigest()])
    for s, idx in sel_idx.items():
        fp['selectors'][s.easier_hint_name] = [int(idx.shape[0]), _sha(idx)]
    for r, idx in red_idx.items():
        fp['reducers'][r.easier_hint_name] = [int(r.n), int(idx.shape[0]), _sha(idx)]
    for o in objs:
        if isinstance(o, _M.Tensor):
            fp['tensors'][o.easier_hint_name] = [list(o.easier_data_loader.shape),
                                                 bool(o.is_partition)]
    return fp


def load(load_dir: str) -> Optional[dict]:
    """The dump under load_dir: {'fingerprint', 'orders': {space name:
    (kind, order)}}, or None (logged) when there is none or it cannot be read."""
    jit_dir = os.path.join(os.path.expanduser(load_dir), 'jit')
    jpath, npath = os.path.join(jit_dir, JSON_NAME), os.path.join(jit_dir, NPZ_NAME)
    if not (os.path.exists(jpath) and os.path.exists(npath)):
        logger.info(f'no compile cache in {jit_dir}; compiling from scratch')
        return None
    try:
        with open(jpath) as f:
            meta = json.load(f)
        with np.load(npath, allow_pickle=False) as z:
            orders = {name: (meta['kinds'][k], z[f'order_{k}'].astype(np.int64))
                      for k, name in enumerate(meta['spaces'])}
        return {'fingerprint': meta['fingerprint'], 'orders': orders}
    except Exception as ex:   # a corrupted dump: the reference recompiles too
        logger.info(f'compile cache in {jit_dir} unreadable ({ex}); compiling from scratch')
        return None


def matching_orders(cache: Optional[dict], fp: dict, names: Dict[int, str],
                    space_len: Dict[int, int]) -> Optional[Dict[int, np.ndarray]]:
    """The cached (kind, order) keyed by space root when the dump matches
    the program (fingerprint, every space's length, permutations), else None."""
    if cache is None:
        return None
    if cache['fingerprint'] != fp:
        logger.info('compile cache does not match the program (changed graph, '
                    'indices, shapes or reorder mode); compiling from scratch')
        return None
    by_name = {v: k for k, v in names.items()}
    out = {}
    for name, (kind, order) in cache['orders'].items():
        root = by_name.get(name)
        if root is None or kind not in ('rcm', 'edge') or \
                order.shape[0] != space_len.get(root, order.shape[0]) or \
                not np.array_equal(np.sort(order), np.arange(order.shape[0])):
            logger.info(f'compile cache space {name!r} does not fit; compiling from scratch')
            return None
        out[root] = (kind, order)
    return out


def dump(modules: List[_M.Module], dump_dir: str) -> None:
    """Write the compile cache of compiled modules to dump_dir/jit/ (the whole
    program they were compiled in). esr.compile(modules, load_dir=dump_dir)
    reuses it for the same program."""
    progs = {id(getattr(m, 'easier_program', None)): getattr(m, 'easier_program', None)
             for m in modules}
    if None in progs.values() or not progs:
        raise EasierJitException('dump() needs modules compiled with esr.compile()')
    if len(progs) != 1:
        raise EasierJitException('dump() takes modules of one esr.compile() call')
    prog = next(iter(progs.values()))
    jit_dir = os.path.join(os.path.expanduser(dump_dir), 'jit')
    os.makedirs(jit_dir, exist_ok=True)
    names = prog.space_name
    spaces = sorted(names[root] for root in prog.space_order)
    by_name = {names[root]: root for root in prog.space_order}
    arrays = {f'order_{k}': np.asarray(prog.space_order[by_name[n]], np.int64)
              for k, n in enumerate(spaces)}
    kinds = [prog.space_kind[by_name[n]] for n in spaces]
    np.savez(os.path.join(jit_dir, NPZ_NAME), **arrays)
    with open(os.path.join(jit_dir, JSON_NAME), 'w') as f:
        json.dump({'fingerprint': prog.fingerprint, 'spaces': spaces, 'kinds': kinds}, f)
    logger.info(f'compile cache of {len(spaces)} renumbered spaces written to {jit_dir}')
