"""Ahead-of-time passes of esr.compile(..., backend='hip') for one GPU.

What the reference does in easier/core/jit.py:294-419 with several passes
(collective_initialization, tensor_grouping, reducer_binding,
tensor_group_partition, sparse_encoding, dataflow_distribution) collapses,
for a single partition, to:

1. trace every esr.Module (same leaf rules as jit.py:78-121) and validate the
   program (collective_initialization.py:20-106);
2. group tensors into index spaces by dataflow (tensor_grouping.py:62-332);
3. renumber: each vertex/cell space that some Reducer writes and some
   Selector reads over a shared edge space gets a reverse Cuthill-McKee order
   of the graph those index pairs form; each edge space is stable-sorted by
   its bound Reducer's (renumbered) index -- at world size 1 the reference
   likewise sorts edges stably by the bound Reducer
   (sparse_encoding.py:244-338, passes/utils.py:301-336), which fixes the
   per-row summation order to the original edge order;
4. load every tensor in the new order, rewrite Selector/Reducer indices as
   int32, build each Reducer's CSR (rowptr, optional permutation for a
   Reducer that does not fix its edge space's order, and the list of rows
   longer than the sequential limit) and move everything to the GPU.
"""
import ctypes
import math
import operator
from collections import OrderedDict
from typing import Dict, List, Optional

import numpy as np
import torch
from torch import fx, nn
from torch.fx import Tracer
from torch.fx.proxy import Proxy

from .. import _native as nat
from . import module as _M
from .utils import EasierJitException, logger


# ---------------------------------------------------------------------------
# tracing
# ---------------------------------------------------------------------------
class EasierProxy(Proxy):
    def __getattr__(self, name):
        if name in ('shape', 'size', 'dim', 'ndim', 'numel', 'nelement'):
            raise NotImplementedError(
                "Currently EASIER does not support accessing common"
                f" hyperparameter information '{name}' during tracing."
                " If needed, those values must be calculated ahead-of-time.")
        return super().__getattr__(name)

    def __setitem__(self, indices, value):
        return self.tracer.create_proxy(
            'call_function', operator.setitem, (self, indices, value), {})


# Captured at import: while tracing, the autowrapped module's names are
# replaced by FX wrappers, so the classes must not be looked up there.
_LEAF_TYPES = (_M.Module, _M.Selector, _M.Reducer)
_TENSOR_TYPE = _M.Tensor


class EasierTracer(Tracer):
    """esr.Module / Selector / Reducer are leaves; esr aggregators and math
    functions are recorded as calls."""

    def __init__(self):
        import easier_amd
        super().__init__(autowrap_modules=(math, easier_amd, _M))

    def is_leaf_module(self, m: nn.Module, qualname: str) -> bool:
        if isinstance(m, _LEAF_TYPES):
            return True
        return super().is_leaf_module(m, qualname)

    def proxy(self, node):
        return EasierProxy(node, self)

    def create_arg(self, a):
        # An esr.Tensor a forward reaches through a closure rather than an
        # attribute (M = lambda r: r / A.Av passed to a solver whose module
        # does not own A) is registered on the traced module, so the graph
        # names it by an attribute path like every other tensor.
        if isinstance(a, _TENSOR_TYPE) and \
                all(p is not a for _, p in self.root.named_parameters(remove_duplicate=False)):
            self.root.register_parameter(f'easier_captured_{len(self.root._parameters)}', a)
        return super().create_arg(a)


def get_easier_objects(top_modules) -> 'OrderedDict':
    """Every esr.Module/Selector/Reducer/Tensor reachable from the roots,
    with the attribute paths naming it (first path = hint name)."""
    objs: 'OrderedDict' = OrderedDict()
    for i, root in enumerate(top_modules):
        base = f'({root.__class__.__name__}#{i})'
        for name, sub in root.named_modules(prefix=base, remove_duplicate=False):
            if isinstance(sub, (_M.Module, _M.Selector, _M.Reducer)):
                objs.setdefault(sub, []).append(name)
            for pname, p in sub.named_parameters(prefix=name, recurse=False,
                                                 remove_duplicate=False):
                if isinstance(p, _M.Tensor):
                    objs.setdefault(p, []).append(pname)
    return objs


def _check_syntax(module: _M.Module, graph: fx.Graph):
    """collective_initialization.py:20-73."""
    for node in graph.nodes:
        if node.op == 'placeholder':
            raise EasierJitException(
                "easier.Module.forward() cannot have parameters")
        if node.op == 'call_function':
            fn = node.target
            mod = getattr(fn, '__module__', '') or ''
            if mod.startswith('easier_amd') and fn not in _M.easier_aggregators:
                raise EasierJitException(f"Unexpected function call to {fn}")
        if node.op == 'call_module':
            sub = module.get_submodule(node.target)
            if isinstance(sub, _M.Module):
                if len(node.users) > 0:
                    raise EasierJitException(
                        "The result of easier.Module.forward() is always"
                        " None and cannot be used")
                if len(node.all_input_nodes) > 0:
                    raise EasierJitException(
                        "easier.Module.forward() cannot have parameters")
            elif not isinstance(sub, (_M.Selector, _M.Reducer)):
                raise EasierJitException(
                    f"torch.nn module {sub} is not supported, consider using"
                    " torch.nn.functional function instead")
        if node.op == 'output' and node.args != (None,):
            raise EasierJitException(
                "easier.Module.forward() cannot have return value")


def resolve_attr(module: nn.Module, target: str):
    obj = module
    for part in target.split('.'):
        obj = getattr(obj, part)
    return obj


# ---------------------------------------------------------------------------
# index-space grouping
# ---------------------------------------------------------------------------
class _UF:
    def __init__(self):
        self.parent: List[int] = []

    def new(self) -> int:
        self.parent.append(len(self.parent))
        return len(self.parent) - 1

    def find(self, x: int) -> int:
        while self.parent[x] != x:
            self.parent[x] = self.parent[self.parent[x]]
            x = self.parent[x]
        return x

    def union(self, a: int, b: int):
        ra, rb = self.find(a), self.find(b)
        if ra != rb:
            self.parent[max(ra, rb)] = min(ra, rb)


_FULL_REDUCE_METHODS = {'sum', 'norm', 'prod', 'amax', 'amin', 'max', 'min',
                        'mean', 'item'}


class SpaceAnalysis:
    """Union-find of the first-dimension index spaces of every value
    (the role of passes/tensor_grouping.py)."""

    def __init__(self):
        self.uf = _UF()
        self.tensor_var: Dict[_M.Tensor, int] = {}
        self.sel_vars: Dict[_M.Selector, tuple] = {}
        self.red_vars: Dict[_M.Reducer, tuple] = {}
        self.node_var: Dict[fx.Node, Optional[int]] = {}
        self.lengths: Dict[int, set] = {}

    def _var_of_tensor(self, t: _M.Tensor) -> int:
        if t not in self.tensor_var:
            v = self.uf.new()
            self.tensor_var[t] = v
            self._len(v, t.easier_data_loader.shape[0])
        return self.tensor_var[t]

    def _len(self, v, n):
        self.lengths.setdefault(v, set()).add(int(n))

    def selector_vars(self, s: _M.Selector):
        if s not in self.sel_vars:
            vin, vout = self.uf.new(), self.uf.new()
            self._len(vout, s.easier_data_loader.shape[0])
            self.sel_vars[s] = (vin, vout)
        return self.sel_vars[s]

    def reducer_vars(self, r: _M.Reducer):
        if r not in self.red_vars:
            vin, vout = self.uf.new(), self.uf.new()
            self._len(vin, r.easier_data_loader.shape[0])
            self._len(vout, r.n)
            self.red_vars[r] = (vin, vout)
        return self.red_vars[r]

    def run(self, module: _M.Module, graph: fx.Graph):
        for t in module.parameters():
            if isinstance(t, _M.Tensor) and t.is_partition:
                self._var_of_tensor(t)
        for node in graph.nodes:
            self.node_var[node] = self._visit(module, node)

    def _args_vars(self, node):
        vs = []

        def f(a):
            if isinstance(a, fx.Node):
                v = self.node_var.get(a)
                if v is not None:
                    vs.append(v)
            return a
        fx.node.map_arg(node.args, f)
        fx.node.map_arg(node.kwargs, f)
        return vs

    def _visit(self, module, node) -> Optional[int]:
        if node.op == 'get_attr':
            obj = resolve_attr(module, node.target)
            if isinstance(obj, _M.Tensor) and obj.is_partition:
                return self._var_of_tensor(obj)
            return None
        if node.op == 'call_module':
            sub = module.get_submodule(node.target)
            if isinstance(sub, _M.Selector):
                vin, vout = self.selector_vars(sub)
                for v in self._args_vars(node):
                    self.uf.union(v, vin)
                return vout
            if isinstance(sub, _M.Reducer):
                vin, vout = self.reducer_vars(sub)
                x = node.args[0]
                if isinstance(x, fx.Node) and self.node_var.get(x) is not None:
                    self.uf.union(self.node_var[x], vin)
                o = node.kwargs.get('out')
                if isinstance(o, fx.Node) and self.node_var.get(o) is not None:
                    self.uf.union(self.node_var[o], vout)
                return vout
            return None
        if node.op in ('call_function', 'call_method'):
            if node.op == 'call_function' and node.target in _M.easier_aggregators:
                return None
            if node.op == 'call_method' and node.target in _FULL_REDUCE_METHODS:
                dim = node.kwargs.get('dim', node.args[1] if len(node.args) > 1
                                      and node.target != 'norm' else None)
                if node.target == 'norm':
                    dim = node.kwargs.get('dim', node.args[2] if len(node.args) > 2 else None)
                if dim is None or dim == 0:
                    return None
            if node.op == 'call_function' and node.target in (
                    torch.sum, torch.norm, torch.prod, torch.amax, torch.amin,
                    torch.mean):
                dim = node.kwargs.get('dim', node.args[1] if len(node.args) > 1 else None)
                if node.target is torch.norm:
                    dim = node.kwargs.get('dim', node.args[2] if len(node.args) > 2 else None)
                if dim is None or dim == 0:
                    return None
            vs = self._args_vars(node)
            if not vs:
                return None
            for v in vs[1:]:
                self.uf.union(vs[0], v)
            if node.op == 'call_function' and node.target is operator.setitem:
                return None
            return vs[0]
        return None

    def root(self, v):
        return self.uf.find(v)

    def space_lengths(self):
        out: Dict[int, set] = {}
        for v, ls in self.lengths.items():
            out.setdefault(self.uf.find(v), set()).update(ls)
        return out


# ---------------------------------------------------------------------------
# native helpers
# ---------------------------------------------------------------------------
def _i64(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int64)


def csr_from_idx(idx: np.ndarray, n: int):
    idx = _i64(idx)
    m = idx.shape[0]
    rowptr = np.empty(n + 1, dtype=np.int64)
    perm = np.empty(m, dtype=np.int64)
    nat.call('esr_csr_from_idx', idx.ctypes.data if m else None, m, n,
             rowptr.ctypes.data, perm.ctypes.data if m else None)
    return rowptr, perm


def rcm_order(n: int, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a, b = _i64(a), _i64(b)
    order = np.empty(n, dtype=np.int64)
    nat.call('esr_rcm', n, a.ctypes.data if a.size else None,
             b.ctypes.data if b.size else None, a.shape[0], order.ctypes.data)
    return order


# rows per tile of the row kernels (their workgroup size) for reorder='blocks'
BLOCK_ROWS = 256


def block_order(n: int, a: np.ndarray, b: np.ndarray, block: int) -> np.ndarray:
    a, b = _i64(a), _i64(b)
    order = np.empty(n, dtype=np.int64)
    nat.call('esr_block_order', n, a.ctypes.data if a.size else None,
             b.ctypes.data if b.size else None, a.shape[0], int(block), order.ctypes.data)
    return order


def bandwidth(a: np.ndarray, b: np.ndarray) -> int:
    a, b = _i64(a), _i64(b)
    bw = ctypes.c_int64(0)
    nat.call('esr_bandwidth', a.ctypes.data if a.size else None,
             b.ctypes.data if b.size else None, a.shape[0], ctypes.byref(bw))
    return bw.value


def _inverse(order: np.ndarray) -> np.ndarray:
    inv = np.empty_like(order)
    inv[order] = np.arange(order.shape[0], dtype=order.dtype)
    return inv


# ---------------------------------------------------------------------------
# the whole pipeline
# ---------------------------------------------------------------------------
class CompiledProgram:
    """Everything one esr.compile() call produced."""

    def __init__(self):
        self.modules: List[_M.Module] = []
        self.graphs: List[fx.Graph] = []
        self.space_order: Dict[int, Optional[np.ndarray]] = {}
        self.space_len: Dict[int, int] = {}
        self.spaces: Optional[SpaceAnalysis] = None
        self.device: Optional[torch.device] = None
        self.bound_reducer: Dict[int, _M.Reducer] = {}
        self.freeze = False


def _validate_idx(name, idx: np.ndarray, upper: Optional[int], kind: str):
    """collective_initialization.py:76-106."""
    if not np.issubdtype(idx.dtype, np.integer):
        raise TypeError(f"{name}.idx must be integer")
    if idx.ndim != 1:
        raise EasierJitException(f'{name}.idx must be 1-dimensional')
    if idx.shape[0] < 1:
        raise EasierJitException(f'{name}.idx must have shape[0] >= 1')
    lo, hi = int(idx.min()), int(idx.max())
    if lo < 0:
        raise ValueError(f"The minimum of {name}.idx {lo}"
                         " must be greater than or equal 0")
    if upper is not None and not hi < upper:
        what = f'{name}.n {upper}' if kind == 'reducer' else \
            f'the length {upper} of its input'
        raise ValueError(f"The maximum of {name}.idx {hi}"
                         f" must be smaller than {what}")


def compile_program(top_modules: List[_M.Module], reorder: str,
                    device: torch.device, cache: Optional[dict] = None) -> CompiledProgram:
    """cache: a loaded compile cache (core/cache.py); its space orders
    replace the renumbering when it matches this program."""
    from . import cache as _cache
    prog = CompiledProgram()
    prog.device = device
    objs = get_easier_objects(top_modules)
    for obj, names in objs.items():
        obj.easier_hint_name = names[0]
    modules = [o for o in objs if isinstance(o, _M.Module)]
    tracer = EasierTracer()
    graphs = []
    for m in modules:
        g = tracer.trace(m)
        _check_syntax(m, g)
        graphs.append(g)
    prog.modules, prog.graphs = modules, graphs

    sa = SpaceAnalysis()
    for m, g in zip(modules, graphs):
        sa.run(m, g)
    prog.spaces = sa
    selectors = [o for o in objs if isinstance(o, _M.Selector)]
    reducers = [o for o in objs if isinstance(o, _M.Reducer)]
    for s in selectors:
        sa.selector_vars(s)
    for r in reducers:
        sa.reducer_vars(r)
    tensors = [o for o in objs if isinstance(o, _M.Tensor)]

    # space lengths must agree
    lens = sa.space_lengths()
    for root, ls in lens.items():
        if len(ls) != 1:
            raise EasierJitException(
                f'tensors grouped into one index space have different '
                f'lengths {sorted(ls)}')
    space_len = {r: next(iter(ls)) for r, ls in lens.items()}

    for t in tensors:
        if t.is_partition:
            shp = t.easier_data_loader.shape
            if len(shp) < 1 or shp[0] < 1:
                raise EasierJitException(
                    f'Distributed easier.Tensor {t.easier_hint_name}.data '
                    'must have ndim >= 1 and shape[0] >= 1')

    # load indices (host, int64)
    sel_idx = {s: s.easier_data_loader.load_cpu().numpy() for s in selectors}
    red_idx = {r: r.easier_data_loader.load_cpu().numpy() for r in reducers}
    for r, idx in red_idx.items():
        if not isinstance(r.n, int):
            raise TypeError(f"{r.easier_hint_name}.n must be an integer")
        _validate_idx(r.easier_hint_name, idx, r.n, 'reducer')
    for s, idx in sel_idx.items():
        vin = sa.root(sa.sel_vars[s][0])
        _validate_idx(s.easier_hint_name, idx, space_len.get(vin), 'selector')

    prog.fingerprint = _cache.fingerprint(modules, graphs, objs, sel_idx, red_idx, reorder,
                                          nat.lib().esr_abi_version())
    prog.space_name = _cache.space_names(sa, objs)
    cached = _cache.matching_orders(cache, prog.fingerprint, prog.space_name, space_len)
    if cached is not None:
        logger.info(f'compile cache: reusing the order of {len(cached)} spaces')

    # ---- renumbering ------------------------------------------------------
    red_in = {r: sa.root(sa.red_vars[r][0]) for r in reducers}
    red_out = {r: sa.root(sa.red_vars[r][1]) for r in reducers}
    sel_in = {s: sa.root(sa.sel_vars[s][0]) for s in selectors}
    sel_out = {s: sa.root(sa.sel_vars[s][1]) for s in selectors}
    reducer_inputs = set(red_in.values())

    order: Dict[int, Optional[np.ndarray]] = {}
    kind: Dict[int, str] = {}
    # node spaces: RCM over (reducer target, selector source) pairs
    if cached is not None:
        order = {root: o for root, (k, o) in cached.items() if k == 'rcm'}
    elif reorder in ('rcm', 'blocks'):
        pairs: Dict[int, list] = {}
        for r in reducers:
            for s in selectors:
                if sel_out[s] == red_in[r] and sel_in[s] == red_out[r]:
                    pairs.setdefault(red_out[r], []).append(
                        (red_idx[r], sel_idx[s]))
        for space, plist in pairs.items():
            if space in reducer_inputs:
                continue
            a = np.concatenate([p[0] for p in plist])
            b = np.concatenate([p[1] for p in plist])
            n = space_len[space]
            if reorder == 'blocks':
                # tiles of BLOCK_ROWS rows (one workgroup of the row kernels)
                # hold compact pieces of the graph
                o = block_order(n, a, b, BLOCK_ROWS)
                order[space] = o
                inv = _inverse(o)
                logger.debug(f'blocks on a space of {n}: edges within a tile '
                             f'{np.mean(inv[a] // BLOCK_ROWS == inv[b] // BLOCK_ROWS):.3f}')
                continue
            o = rcm_order(n, a, b)
            inv = _inverse(o)
            if bandwidth(inv[a], inv[b]) < bandwidth(a, b):
                order[space] = o
            logger.debug(f'RCM on a space of {n}: bandwidth '
                         f'{bandwidth(a, b)} -> {bandwidth(inv[a], inv[b])}')
    elif reorder != 'none':
        raise EasierJitException(
            f"reorder must be 'rcm', 'blocks' or 'none', got {reorder!r}")
    for root in order:
        kind[root] = 'rcm'

    def inv_of(space):
        o = order.get(space)
        return None if o is None else _inverse(o)

    # edge spaces: stable sort by the bound reducer / first selector
    for r in reducers:
        e = red_in[r]
        if e in order or e in prog.bound_reducer:
            continue
        prog.bound_reducer[e] = r
        if cached is not None:
            if e in cached:
                order[e], kind[e] = cached[e][1], 'edge'
            continue
        inv = inv_of(red_out[r])
        key = red_idx[r] if inv is None else inv[red_idx[r]]
        _, perm = csr_from_idx(key, r.n)
        if not np.array_equal(perm, np.arange(perm.shape[0])):
            order[e], kind[e] = perm, 'edge'
    for s in selectors:
        e = sel_out[s]
        if e in order or e in reducer_inputs or e in prog.bound_reducer:
            continue
        if cached is not None:
            if e in cached:
                order[e], kind[e] = cached[e][1], 'edge'
            prog.bound_reducer[e] = None
            continue
        # a Selector never applied in a traced forward has no known input
        # length: bound its index range by its own maximum
        n_in = space_len.get(sel_in[s], int(sel_idx[s].max()) + 1)
        inv = inv_of(sel_in[s])
        key = sel_idx[s] if inv is None else inv[sel_idx[s]]
        _, perm = csr_from_idx(key, n_in)
        if not np.array_equal(perm, np.arange(perm.shape[0])):
            order[e], kind[e] = perm, 'edge'
        prog.bound_reducer[e] = None
    prog.space_order = order
    prog.space_kind = kind
    prog.space_len = space_len

    # ---- load data in the new order ---------------------------------------
    def to_dev_idx(a: np.ndarray) -> torch.Tensor:
        if a.size and int(a.max()) >= 2 ** 31:
            raise EasierJitException('index exceeds int32 range')
        return torch.from_numpy(a.astype(np.int32)).to(device)

    for s in selectors:
        idx = sel_idx[s]
        o = order.get(sel_out[s])
        if o is not None:
            idx = idx[o]
        inv = inv_of(sel_in[s])
        if inv is not None:
            idx = inv[idx]
        s.idx = to_dev_idx(idx)
        s.easier_idx_max = int(idx.max())
        s.easier_index_status = 'rewritten'
        s.easier_host_idx = idx

    for r in reducers:
        idx = red_idx[r]
        o = order.get(red_in[r])
        if o is not None:
            idx = idx[o]
        inv = inv_of(red_out[r])
        if inv is not None:
            idx = inv[idx]
        rowptr, perm = csr_from_idx(idx, r.n)
        counts = np.diff(rowptr)
        is_sorted = np.array_equal(perm, np.arange(perm.shape[0]))
        r.idx = to_dev_idx(idx)
        r.easier_rowptr = to_dev_idx(rowptr)
        r.easier_perm = None if is_sorted else to_dev_idx(perm)
        long_rows = np.nonzero(counts > nat.SEQ_ROW_MAX)[0]
        r.easier_long_rows = to_dev_idx(long_rows) if long_rows.size else None
        r.easier_index_status = 'rewritten'
        r.easier_host_idx = idx
        r.easier_host_rowptr = rowptr
        r.easier_is_full = bool(np.all(counts > 0))

    for t in tensors:
        data = t.easier_data_loader.load_cpu()
        if t.is_partition:
            space = sa.root(sa._var_of_tensor(t))
            o = order.get(space)
            if o is not None:
                data = data[torch.from_numpy(o)]
                t.easier_order = torch.from_numpy(o)
            else:
                t.easier_order = None
        t.data = data.contiguous().to(device)
        t.easier_data_ready = True
    return prog
