"""Fused lowering of esr.Module forwards into row kernels.

The reference evaluates a forward node by node (runtime/jit_engine/
jit_engine.py:1114-1172): every Selector materialises an edge-sized tensor,
every elementwise op another, every Reducer a scatter. Here a forward is
symbolically executed once into a value graph and scheduled into *row
kernels*: a kernel iterates the rows of one index space (cells, vertices) and

  - evaluates every Reducer writing that space as a loop over the row's CSR
    range (edges sorted by row, so the loop is the reference's sequential
    per-row sum in edge order), with the edge expression inlined: edge
    tensors are read at the edge position, Selector outputs are read through
    the Selector's index (or at the row itself when the Selector's index is
    the Reducer's row index),
  - evaluates the elementwise epilogue of the row,
  - writes only what a later kernel, a Selector, or an assignment to an
    esr.Tensor needs.

A value gathered through a Selector must be complete before the gathering
kernel starts, so gathers bump the consumer to the next stage; everything
else fuses. An assignment to a tensor is placed after every kernel that reads
the old contents at other rows. The kernels of one forward are generated as
C++ from the template below, compiled once with hipRTC for the GPU, launched
in stage order and captured as one HIP graph.

Forwards using anything outside this subset (aggregators, views, replicated
tensor updates, nested modules, multi-column tensors, ...) keep the
node-by-node HipEngine.
"""
import ctypes
import hashlib
import operator
import struct
from typing import Dict, List, Optional

import numpy as np
import torch
from torch import fx

from .. import _native as nat
from . import module as _M
from . import compiler as _compiler
from .compiler import CompiledProgram, resolve_attr
from .utils import logger


class Unsupported(Exception):
    pass


# Code-generation options of the row kernels (all keep the arithmetic and its
# order unchanged):
#   stage_edges  - a dense Reducer group (at most 8 edges per row) first copies
#                  the block's contiguous edge range of every edge array it
#                  reads into LDS with coalesced loads, then reads the rows'
#                  edges from LDS;
#   nt_loads     - edge-array loads carry the nontemporal hint (streamed once);
#   nt_stores    - row stores carry the nontemporal hint;
#   block        - rows (threads) per workgroup;
#   xcd_remap    - consecutive row tiles are placed on one XCD (its L2 then
#                  holds the rows their neighbours gather); off, the 8 XCDs
#                  stream through neighbouring rows together, which keeps the
#                  ~20 concurrent array streams of a stage kernel on fewer
#                  DRAM pages (measured faster for the SWE stage kernels).
#   sparse_rows  - a Reducer with few non-empty rows (at most 1 in 8) finds a
#                  row's entries through a bitmap of the non-empty rows (one
#                  64-bit word and one prefix count per 64 rows) and the row
#                  pointers of the non-empty rows only, instead of reading
#                  two row pointers for every row.
# Defaults are the fastest measured combination for the 2048 x 2048 SWE step.
#   idx16        - in dense Reducer groups, a Selector's index is read as int16
#                  offsets from the edge's row when all of them fit (a mesh
#                  renumbered for locality), half the bytes of int32 indices.
#   xcd_group    - G > 0: consecutive groups of G tiles per XCD (esr_group_remap);
#                  overrides xcd_remap. Measured on the 2048 x 2048 SWE step
#                  (bytes fetched from beyond L2 / algorithmic, ms per step):
#                  no remap 1.285 / 1.189, xcd_remap 1.002 / 1.217,
#                  G = 16 1.105 / 1.164, G = 64 1.035 / 1.175, G = 256
#                  1.009 / 1.196; measured again with the current kernels
#                  (ms per step, two runs each): G = 8 1.163, G = 16
#                  1.161 / 1.167, G = 32 1.179 / 1.163, G = 64 1.177 / 1.180.
#   rows_per_thread - 2: a thread evaluates rows r0 and r0 + 1 (r0 even); the
#                  row arrays are read and written as 2-element vectors (16
#                  bytes in float64), the edge loops stay per row. Measured on
#                  the 2048 x 2048 SWE step: 1.38 vs 1.18 ms (half the threads
#                  leave fewer gathers in flight than the wider row accesses
#                  save), so the default is 1.
#   pack_intermediates - the values one kernel materialises for later kernels
#                  share one buffer, interleaved by row (value j of row r at
#                  r * width + j): one array stream instead of `width` for
#                  their writer and readers (rows_per_thread 1 only).
#                  Measured on the 2048 x 2048 SWE step: 1.41 vs 1.18 ms (a
#                  load instruction then spans width x the rows of one
#                  stream and uses 1/width of each line it touches), so the
#                  default is off.
#   waves_per_eu - N > 0: ask the register allocator for N waves per SIMD
#                  (amdgpu_waves_per_eu); 0 leaves occupancy to the compiler.
#                  On the SWE 2048^2 float64 step the four stage kernels run
#                  8/7/7/6 waves by default; N = 8 left the first three at
#                  256/318/318 us and made the fourth spill 44 bytes per lane
#                  (292 -> 417 us, step 1.18 -> 1.31 ms), so the default is 0.
#   remat        - an elementwise value that later kernels only gather (an RK
#                  stage input h + c * d) is not written to a buffer: each
#                  gather evaluates it at the source row from the arrays it
#                  is made of (at most two, each stored anyway: tensors not
#                  reassigned before the gather, and values other kernels
#                  read), so one array write and one array read become reads
#                  of arrays the gathering kernel mostly reads at its own rows.
#   late_eval    - an elementwise value whose users all run in later kernels
#                  is evaluated in the first of them instead of its earliest
#                  kernel when that lowers the forward's algorithmic bytes
#                  (e.g. d1 + d2 in the kernel that reads d2 anyway, saving
#                  the partial sum's write and read); greedy, program order.
#   dict_max     - with esr.compile(..., freeze=True): a partition tensor no
#                  forward of the program writes, whose values take at most
#                  dict_max distinct bit patterns, is read through a
#                  dictionary: one uint8 code per row (shared by all such
#                  tensors of one index space when their tuples of values take
#                  at most dict_max patterns together) and a table of values
#                  staged in LDS; a tensor with one value becomes a kernel
#                  constant read once. Bit-exact (the table holds the
#                  tensor's own values); tensors under dict_min_rows rows are
#                  left alone.
#   ell          - with freeze=True: a dense Reducer group (at most 8 edges per
#                  row, no permutation) reads its edge data from padded
#                  per-row records (ELL: S = 4 or 8 slots per row, slot k of
#                  row r at r S + k, empty slots after the row's edges) built
#                  at compile time -- the int16 source offsets, the dictionary
#                  codes and any other edge array -- each loaded with one
#                  vector load per row, and the row's edge count (uint8)
#                  instead of two row pointers: no row-pointer load ahead of
#                  the edge loads, one instruction per array per row. Slots
#                  fold in the same order, empty slots are skipped.
#   halo         - with ell: the values a group gathers through a Selector
#                  are staged in LDS per tile of rows -- the tile's own rows
#                  (each thread its row) and the tile's halo (the other source
#                  rows its edges read, a list built at compile time, at most
#                  halo_max per tile) -- and each edge slot reads LDS at a
#                  uint16 index; one coalesced load and one halo load per
#                  array and tile replace a gather per edge slot (the
#                  gathers' address processing bounds the kernels otherwise).
#                  Tiles whose halo does not fit keep the gathers.
OPTIONS = {'stage_edges': False, 'nt_loads': False, 'nt_stores': True, 'block': 256,
           'xcd_remap': False, 'sparse_rows': True, 'idx16': True, 'xcd_group': 16,
           'rows_per_thread': 1, 'pack_intermediates': False, 'waves_per_eu': 0,
           'remat': True, 'late_eval': True, 'dict_max': 256, 'dict_min_rows': 4096,
           'ell': True, 'dict_lds': True, 'persist': 0, 'halo': True, 'halo_max': 1024}
_LDS_BUDGET = 48 * 1024


def _delta16(sel, red):
    """int16 device array idx[e] - row(e) for the edges of a Reducer-ordered
    edge space (row(e) the Reducer row of edge e), or None when an offset
    does not fit in int16."""
    cache = sel.__dict__.setdefault('easier_delta16', {})
    if id(red) in cache:
        return cache[id(red)]
    hp = np.asarray(red.easier_host_rowptr, np.int64)
    idx = np.asarray(sel.easier_host_idx, np.int64)
    rows = np.repeat(np.arange(hp.shape[0] - 1, dtype=np.int64), np.diff(hp))
    out = None
    if rows.shape == idx.shape:
        d = idx - rows
        if d.size == 0 or (d.min() >= -32768 and d.max() <= 32767):
            out = torch.from_numpy(d.astype(np.int16)).to(red.easier_rowptr.device)
    cache[id(red)] = out
    return out


def _sparse_rows(red):
    """(bitmap, prefix, compact row pointers) of a Reducer's non-empty rows,
    built once from its host CSR: bit r % 64 of word r // 64 is set when row
    r has entries; prefix[w] counts the set bits of words before w; the c-th
    non-empty row's entries are [crow[c], crow[c + 1])."""
    cached = getattr(red, 'easier_sparse_rows', None)
    if cached is not None:
        return cached
    hp = np.asarray(red.easier_host_rowptr, np.int64)
    n = hp.shape[0] - 1
    nonempty = np.nonzero(hp[1:] > hp[:-1])[0]
    nw = (n + 63) // 64
    bits = np.zeros((nw, 64), np.uint64)
    bits.reshape(-1)[nonempty] = 1
    words = (bits << np.arange(64, dtype=np.uint64)).sum(1, dtype=np.uint64)
    prefix = np.concatenate([[0], np.cumsum(bits.sum(1, dtype=np.int64))[:-1]])
    crow = np.concatenate([hp[nonempty], [hp[-1]]])
    dev = red.easier_rowptr.device
    out = (torch.from_numpy(words.view(np.int64)).to(dev),
           torch.from_numpy(prefix.astype(np.int32)).to(dev),
           torch.from_numpy(crow.astype(np.int32)).to(dev))
    red.easier_sparse_rows = out
    return out


class _Enc:
    """A frozen tensor read through a dictionary: codes (uint8 per row, or
    None when the tensor holds one value) index lut (its distinct values)."""

    def __init__(self, codes, lut):
        self.codes, self.lut = codes, lut
        self.k = int(lut.numel())


def _written_tensors(prog: CompiledProgram) -> set:
    """ids of the esr.Tensors some forward of the program may write: assigned
    (setitem), updated in place (a method or function ending in '_'), an
    `out=` target, or handed to a module call other than a Selector or
    Reducer (through views taken with getitem)."""
    out = set()
    for m, g in zip(prog.modules, prog.graphs):
        env = {}
        for node in g.nodes:
            if node.op == 'get_attr':
                try:
                    obj = resolve_attr(m, node.target)
                except Exception:
                    continue
                if isinstance(obj, _M.Tensor):
                    env[node] = obj

        def mark(a, depth=0):
            if not isinstance(a, fx.Node) or depth > 16:
                return
            if a in env:
                out.add(id(env[a]))
            elif a.op in ('call_function', 'call_method') and a.args and (
                    a.target is operator.getitem or
                    a.target in ('view', 'reshape', 'view_as', 'expand', 'squeeze',
                                 'unsqueeze', 'flatten', 'narrow', 'select', '__getitem__')):
                mark(a.args[0], depth + 1)
        for node in g.nodes:
            name = node.target if node.op == 'call_method' else \
                getattr(node.target, '__name__', '')
            if node.op == 'call_function' and node.target is operator.setitem:
                mark(node.args[0])
            elif node.op in ('call_method', 'call_function') and isinstance(name, str) \
                    and name.endswith('_') and not name.endswith('__') and node.args:
                mark(node.args[0])
            if 'out' in node.kwargs:
                mark(node.kwargs['out'])
            if node.op == 'call_module':
                sub = m.get_submodule(node.target)
                if not isinstance(sub, (_M.Selector, _M.Reducer)):
                    for a in node.args:
                        mark(a)
    return out


def _encode_group(ts: list, prog: CompiledProgram) -> Dict[int, _Enc]:
    """Dictionaries of the frozen tensors `ts` of one index space (once per
    compile): tensors with one value become constants; the others share one
    code array when their value tuples take at most dict_max patterns, else
    each tensor with at most dict_max values gets its own."""
    cache = prog.__dict__.setdefault('easier_dict_cache', {})
    key = tuple(id(t) for t in ts)
    if key in cache:
        return cache[key]
    maxk = int(OPTIONS['dict_max'])
    ftype = ts[0].dtype
    itype = np.int64 if ftype == torch.float64 else np.int32
    dev = ts[0].data.device
    per = []
    for t in ts:
        h = t.data.detach().cpu().numpy().view(itype)
        if np.unique(h[::max(1, h.size // 65536)]).size > maxk:
            per.append(None)
            continue
        u, inv = np.unique(h, return_inverse=True)
        per.append(None if u.size > maxk else (u, inv.reshape(-1)))

    def lut_of(vals):
        return torch.from_numpy(np.ascontiguousarray(vals).view(
            np.float64 if ftype == torch.float64 else np.float32).copy()).to(dev)
    res: Dict[int, _Enc] = {}
    multi = []
    for i, x in enumerate(per):
        if x is None:
            continue
        if x[0].size == 1:
            res[id(ts[i])] = _Enc(None, lut_of(x[0]))
        else:
            multi.append(i)
    multi = multi[:7]                      # 256^7 patterns fit an int64 key
    joint = False
    if len(multi) >= 2:
        jk = np.zeros(per[multi[0]][1].shape[0], np.int64)
        for i in multi:
            jk = jk * maxk + per[i][1]
        ju, jinv = np.unique(jk, return_inverse=True)
        if ju.size <= maxk:
            codes = torch.from_numpy(jinv.reshape(-1).astype(np.uint8)).to(dev)
            rem = ju.copy()
            for i in reversed(multi):
                res[id(ts[i])] = _Enc(codes, lut_of(per[i][0][rem % maxk]))
                rem //= maxk
            joint = True
    if not joint:
        for i in multi:
            u, inv = per[i]
            res[id(ts[i])] = _Enc(torch.from_numpy(inv.astype(np.uint8)).to(dev), lut_of(u))
    cache[key] = res
    return res


def _frozen_encoding(p: 'Program', prog: CompiledProgram) -> Dict[int, _Enc]:
    """OPTIONS['dict_max'] under esr.compile(..., freeze=True): dictionaries
    of the program's constant partition tensors, by id."""
    written = _written_tensors(prog)
    groups: Dict[object, list] = {}
    for v in p.values:
        if v.kind != 'leaf' or id(v.obj) in written:
            continue
        t = v.obj
        if not t.is_partition or t.data.dim() != 1 or \
                t.dtype not in (torch.float32, torch.float64) or \
                t.data.numel() < int(OPTIONS['dict_min_rows']):
            continue
        lst = groups.setdefault(v.space, [])
        if all(x is not t for x in lst):
            lst.append(t)
    out: Dict[int, _Enc] = {}
    for ts in groups.values():
        out.update(_encode_group(ts, prog))
    return out


def _ell_array(prog: CompiledProgram, red, S: int, key, host_values: np.ndarray, dev):
    """An edge array of a Reducer-ordered edge space in the padded per-row
    layout (row r's k-th edge at r * S + k, zeros in empty slots), once per
    compile."""
    cache = prog.__dict__.setdefault('easier_ell_cache', {})
    ck = ('arr', key, id(red), S)
    if ck not in cache:
        slot = _ell_slots(prog, red, S)
        out = np.zeros(slot.shape, host_values.dtype)
        ok = slot >= 0
        out[ok] = host_values[slot[ok]]
        cache[ck] = torch.from_numpy(out.reshape(-1)).to(dev)
    return cache[ck]


def _ell_slots(prog: CompiledProgram, red, S: int) -> np.ndarray:
    """(n, S) edge positions of each row's slots, -1 when empty."""
    cache = prog.__dict__.setdefault('easier_ell_cache', {})
    ck = ('slots', id(red), S)
    if ck not in cache:
        hp = np.asarray(red.easier_host_rowptr, np.int64)
        n = hp.shape[0] - 1
        counts = np.diff(hp)
        ne = int(hp[-1])
        rows = np.repeat(np.arange(n, dtype=np.int64), counts)
        k = np.arange(ne, dtype=np.int64) - np.repeat(hp[:-1], counts)
        slot = np.full((n, S), -1, np.int64)
        slot[rows, k] = np.arange(ne, dtype=np.int64)
        cache[ck] = slot
    return cache[ck]


def _ell_counts(prog: CompiledProgram, red, dev):
    cache = prog.__dict__.setdefault('easier_ell_cache', {})
    ck = ('cnt', id(red))
    if ck not in cache:
        counts = np.diff(np.asarray(red.easier_host_rowptr, np.int64))
        cache[ck] = torch.from_numpy(counts.astype(np.uint8)).to(dev)
    return cache[ck]


def _halo_layout(prog: CompiledProgram, sel, red, S: int, B: int, hmax: int, dev):
    """Per tile of B rows: (local index per slot (uint16, n * S), halo rows
    (int32, ntiles * HS), HS) for the Selector's source rows of a
    Reducer-ordered edge space, or None when some tile reads more than hmax
    rows outside itself. Slot k of row r reads LDS entry i: i < B is row
    tile * B + i, i >= B the tile's halo row i - B."""
    cache = prog.__dict__.setdefault('easier_ell_cache', {})
    ck = ('halo', id(sel), id(red), S, B, hmax)
    if ck in cache:
        return cache[ck]
    slot = _ell_slots(prog, red, S)
    n = slot.shape[0]
    idx = np.asarray(sel.easier_host_idx, np.int64)
    rows = np.repeat(np.arange(n, dtype=np.int64)[:, None], S, axis=1)
    src = np.where(slot >= 0, idx[np.maximum(slot, 0)], rows)
    tile = rows // B
    loc = src - tile * B
    ext = (loc < 0) | (loc >= B)
    ntiles = (n + B - 1) // B
    keys = tile[ext] * n + src[ext]
    uk = np.unique(keys)
    ut = uk // n
    counts = np.bincount(ut, minlength=ntiles)
    hs = int(counts.max()) if counts.size else 0
    if hs > hmax:
        cache[ck] = None
        return None
    hs = max(64, (hs + 63) // 64 * 64)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    pos = np.arange(uk.size) - starts[ut]
    halo = (np.arange(ntiles, dtype=np.int64)[:, None] * B).repeat(hs, 1)
    halo[ut, pos] = uk % n
    halo = np.minimum(halo, n - 1)
    local = loc.copy()
    local[ext] = B + pos[np.searchsorted(uk, keys)]
    out = (torch.from_numpy(local.astype(np.uint16).reshape(-1)).to(dev),
           torch.from_numpy(halo.astype(np.int32).reshape(-1)).to(dev), hs)
    cache[ck] = out
    return out


class _LocalIdx(str):
    """Name of a kernel variable holding an LDS-local gather index (halo
    mode); `sel` identifies the Selector whose staged values it indexes."""

    def __new__(cls, name, sel):
        obj = str.__new__(cls, name)
        obj.sel = sel
        return obj


class _NoHalo(Exception):
    """A kernel's halo staging exceeds the LDS budget."""


_EW_BIN = {
    operator.add: 'add', operator.sub: 'sub', operator.mul: 'mul',
    operator.truediv: 'div', operator.pow: 'pow',
    operator.lt: 'lt', operator.le: 'le', operator.gt: 'gt', operator.ge: 'ge',
    operator.eq: 'eq', operator.ne: 'ne',
    torch.add: 'add', torch.sub: 'sub', torch.mul: 'mul', torch.div: 'div',
    torch.pow: 'pow', torch.maximum: 'maximum', torch.minimum: 'minimum',
}
_EW_UN = {
    operator.neg: 'neg', torch.neg: 'neg', torch.abs: 'abs', operator.abs: 'abs',
    torch.exp: 'exp', torch.sqrt: 'sqrt', torch.sin: 'sin', torch.cos: 'cos',
    torch.tanh: 'tanh', torch.log: 'log', torch.sign: 'sign',
}
_M_BIN = {'add': 'add', 'sub': 'sub', 'mul': 'mul', 'div': 'div', 'pow': 'pow'}
_M_UN = {'neg': 'neg', 'abs': 'abs', 'exp': 'exp', 'sqrt': 'sqrt', 'sin': 'sin',
         'cos': 'cos', 'tanh': 'tanh', 'log': 'log', 'sign': 'sign'}
_M_INPLACE = {'add_': 'add', 'sub_': 'sub', 'mul_': 'mul', 'div_': 'div'}
_PRED = {'lt', 'le', 'gt', 'ge', 'eq', 'ne'}


class V:
    """A value of the fused program."""
    _n = 0

    def __init__(self, kind, space=None, args=(), op=None, obj=None,
                 const=None, dtype=None):
        V._n += 1
        self.id = V._n
        self.kind = kind          # leaf scalar const ew gather reduce store
        self.space = space        # index-space root, or None (scalar/const)
        self.args = list(args)
        self.op = op
        self.obj = obj
        self.const = const
        self.dtype = dtype
        self.stage = -1
        self.users: List['V'] = []
        self.mat = False          # written to a buffer by its kernel
        self.buf: Optional[torch.Tensor] = None
        self.pack = None          # (width, column) in a row-interleaved buffer
        for a in self.args:
            a.users.append(self)


class _Version:
    def __init__(self, value: V):
        self.value = value
        self.reads = []           # (stage, is_gather)


class Program:
    """Symbolic execution of one forward graph."""

    def __init__(self, module: _M.Module, graph: fx.Graph, prog: CompiledProgram):
        self.m, self.g, self.prog = module, graph, prog
        self.sa = prog.spaces
        self.values: List[V] = []
        self.stores: List[V] = []
        self.version: Dict[_M.Tensor, _Version] = {}
        self.leaf_of: Dict[_M.Tensor, V] = {}
        self.env: Dict[fx.Node, object] = {}
        self.dtype = None

    # -- helpers --------------------------------------------------------------
    def _new(self, *a, **k):
        v = V(*a, **k)
        self.values.append(v)
        return v

    def _check_dtype(self, dt):
        if dt not in (torch.float32, torch.float64):
            raise Unsupported(f'dtype {dt}')
        if self.dtype is None:
            self.dtype = dt
        elif self.dtype != dt:
            raise Unsupported('mixed float dtypes')

    def _tensor_value(self, t: _M.Tensor) -> V:
        if t not in self.version:
            if t.is_replica:
                if t.data.numel() != 1:
                    raise Unsupported('replicated tensor with more than one element')
                self._check_dtype(t.dtype)
                v = self._new('scalar', None, obj=t, dtype=t.dtype)
            else:
                if t.data.dim() != 1:
                    raise Unsupported('multi-column tensor')
                self._check_dtype(t.dtype)
                v = self._new('leaf', self.sa.root(self.sa.tensor_var[t]), obj=t,
                              dtype=t.dtype)
            self.leaf_of[t] = v
            self.version[t] = _Version(v)
        return self.version[t].value

    def _arg(self, a):
        if isinstance(a, fx.Node):
            x = self.env[a]
            if isinstance(x, _M.Tensor):
                return self._tensor_value(x)
            return x
        if isinstance(a, bool) or not isinstance(a, (int, float)):
            raise Unsupported(f'argument {a!r}')
        return self._new('const', None, const=float(a))

    def _space_of(self, args):
        sp = {a.space for a in args if a.space is not None}
        if len(sp) > 1:
            raise Unsupported('operands from different index spaces')
        return next(iter(sp)) if sp else None

    def _ew(self, op, args):
        args = [self._arg(a) for a in args]
        if op == 'pow' and args[1].kind != 'const':
            raise Unsupported('pow with a tensor exponent')
        return self._new('ew', self._space_of(args), args=args, op=op,
                         dtype=self.dtype)

    def _store(self, target, value: V):
        if not isinstance(target, _M.Tensor) or not target.is_partition:
            raise Unsupported('assignment to a non-partition tensor')
        old = self._tensor_value(target)
        if value.kind == 'const':
            value = self._new('ew', old.space, args=[value], op='bcast', dtype=self.dtype)
        if value.space not in (None, old.space):
            raise Unsupported('assignment across index spaces')
        st = self._new('store', old.space, args=[value], obj=target, dtype=self.dtype)
        st.old = self.version[target]
        self.stores.append(st)
        self.version[target] = _Version(value)
        return st

    # -- symbolic execution -------------------------------------------------
    def run(self):
        for node in self.g.nodes:
            op, t = node.op, node.target
            if node.op == 'output':
                continue
            if op == 'get_attr':
                obj = resolve_attr(self.m, t)
                if not isinstance(obj, _M.Tensor):
                    raise Unsupported(f'attribute {t}')
                self.env[node] = obj
                continue
            if node.kwargs and not (op == 'call_method' and t in _M_INPLACE
                                    and set(node.kwargs) <= {'alpha'}
                                    and node.kwargs.get('alpha', 1) == 1):
                raise Unsupported(f'keyword arguments in {node.format_node()}')
            if op == 'call_module':
                sub = self.m.get_submodule(t)
                x = self._arg(node.args[0])
                if isinstance(sub, _M.Selector):
                    if x.space != self.sa.root(self.sa.sel_vars[sub][0]):
                        raise Unsupported('selector input space')
                    self.env[node] = self._new(
                        'gather', self.sa.root(self.sa.sel_vars[sub][1]), args=[x],
                        obj=sub, dtype=self.dtype)
                elif isinstance(sub, _M.Reducer):
                    if sub.reduce != 'sum':
                        raise Unsupported('non-sum reducer')
                    if x.space != self.sa.root(self.sa.red_vars[sub][0]):
                        raise Unsupported('reducer input space')
                    self.env[node] = self._new(
                        'reduce', self.sa.root(self.sa.red_vars[sub][1]), args=[x],
                        obj=sub, dtype=self.dtype)
                else:
                    raise Unsupported('nested module call')
                continue
            if op == 'call_function':
                if t in _EW_BIN:
                    self.env[node] = self._ew(_EW_BIN[t], node.args)
                elif t in _EW_UN:
                    self.env[node] = self._ew(_EW_UN[t], node.args)
                elif t is operator.pos:
                    self.env[node] = self._arg(node.args[0])
                elif t is torch.where:
                    c, a, b = (self._arg(x) for x in node.args)
                    self.env[node] = self._new('ew', self._space_of([c, a, b]),
                                               args=[c, a, b], op='where',
                                               dtype=self.dtype)
                elif t is operator.getitem:
                    x, idx = node.args
                    if idx != slice(None) and idx is not Ellipsis:
                        raise Unsupported('indexing')
                    self.env[node] = self._arg(x)
                elif t is operator.setitem:
                    tgt, idx, val = node.args
                    if idx != slice(None) and idx is not Ellipsis:
                        raise Unsupported('partial assignment')
                    self._store(self.env[tgt], self._arg(val))
                    self.env[node] = None
                else:
                    raise Unsupported(f'function {getattr(t, "__name__", t)}')
                continue
            if op == 'call_method':
                if t in _M_INPLACE:
                    tgt = self.env[node.args[0]]
                    val = self._ew(_M_INPLACE[t], node.args)
                    self._store(tgt, val)
                    self.env[node] = self._tensor_value(tgt)
                elif t == 'copy_':
                    tgt = self.env[node.args[0]]
                    self._store(tgt, self._arg(node.args[1]))
                    self.env[node] = self._tensor_value(tgt)
                elif t in _M_BIN:
                    self.env[node] = self._ew(_M_BIN[t], node.args)
                elif t in _M_UN:
                    self.env[node] = self._ew(_M_UN[t], node.args)
                else:
                    raise Unsupported(f'method {t}')
                continue
            raise Unsupported(node.op)
        if not self.stores:
            raise Unsupported('forward assigns nothing')
        if self.dtype is None:
            raise Unsupported('no float tensors')

    # -- scheduling -----------------------------------------------------------
    def schedule(self, late: Optional[Dict[int, int]] = None):
        """Assign every value the stage of the kernel that evaluates it.

        c(leaf) = -1 (ready before the forward); c(gather) = 0 for a leaf
        source, else c(source) + 1 (the source must be complete); c(ew/reduce)
        = max(0, c(args), late[v]) -- `late` holds evaluation stages chosen
        later than the earliest one. Edge-space values are re-evaluated inside each
        reduce loop that uses them, so their evaluation stages are those of
        their consumers. A store runs no earlier than its value and after
        every evaluation that reads the stored tensor's old contents (strictly
        after, for reads at other rows through a Selector)."""
        sa = self.sa
        late = late or {}
        self.edge_spaces = {sa.root(sa.red_vars[r][0]) for r in sa.red_vars}
        for v in self.values:
            if v.kind in ('leaf', 'scalar', 'const'):
                v.stage = -1
            elif v.kind == 'gather':
                src = v.args[0]
                v.stage = 0 if src.kind in ('leaf', 'scalar') else src.stage + 1
            elif v.kind in ('ew', 'reduce'):
                v.stage = max([0, late.get(v.id, 0)] + [a.stage for a in v.args])
        for st in self.stores:
            if st.space in self.edge_spaces:
                raise Unsupported('assignment to an edge tensor')
        # evaluation stages (backwards: users come later in creation order)
        evals: Dict[int, set] = {}
        for v in reversed(self.values):
            if v.kind == 'store':
                continue
            if v.kind in ('gather', 'ew') and v.space in self.edge_spaces:
                s = set()
                for u in v.users:
                    if u.kind == 'reduce':
                        s.add(u.stage)
                    else:
                        s |= evals.get(u.id, {u.stage})
                evals[v.id] = s or {v.stage}
            else:
                evals[v.id] = {v.stage}
        self.evals = evals
        # stores, in program order
        last_store: Dict[_M.Tensor, int] = {}
        for st in self.stores:
            val = st.args[0]
            s = max(0, val.stage, last_store.get(st.obj, 0))
            leaf = self.leaf_of.get(st.obj)
            if leaf is not None:
                for u in leaf.users:
                    if u is st:
                        continue
                    for es in evals.get(u.id, {u.stage}):
                        s = max(s, es + (1 if u.kind == 'gather' else 0))
            st.stage = s
            last_store[st.obj] = s

def _fmt_const(x: float, ctype: str) -> str:
    if x != x:
        return f'({ctype})(__builtin_nan(""))'
    if x in (float('inf'), float('-inf')):
        return f'({ctype})({"-" if x < 0 else ""}__builtin_inf())'
    return f'({ctype})({float(x).hex()})'


_CODE_HEADER = r'''
typedef long long i64;
#define ESR_XCD 8
__device__ __forceinline__ int esr_clampi(int v, int hi) {
  return v < 0 ? 0 : (v > hi ? hi : v);
}
// Workgroup b runs on XCD b % 8 as that XCD's (b / 8)-th: groups of G
// consecutive tiles go to one XCD, XCDs taking consecutive groups, so the 8
// XCDs stream through one window of 8 G tiles together while a tile's
// neighbours mostly stay in its XCD's L2. Tiles past the last whole window
// keep their own index (a bijection on [0, nb)).
__device__ __forceinline__ i64 esr_group_remap(i64 b, i64 nb, i64 G) {
  const i64 full = (nb / (8 * G)) * 8 * G;
  if (b >= full) return b;
  const i64 x = b % 8, s = b / 8;
  return ((s / G) * 8 + x) * G + (s % G);
}
__device__ __forceinline__ i64 esr_remap(i64 b, i64 nb) {
  const i64 q = nb / ESR_XCD, r = nb % ESR_XCD;
  const i64 x = b % ESR_XCD, pos = b / ESR_XCD;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + pos;
}
// two consecutive rows as one vector access (rows r, r + 1; at the last row
// of an odd count, a scalar access, the second lane a copy)
typedef double esr_v2d __attribute__((ext_vector_type(2)));
typedef float esr_v2f __attribute__((ext_vector_type(2)));
__device__ __forceinline__ esr_v2d esr_ld2(const double *p, i64 r, i64 n) {
  if (r + 1 < n) return *reinterpret_cast<const esr_v2d *>(p + r);
  esr_v2d v; v.x = p[r]; v.y = v.x; return v;
}
__device__ __forceinline__ esr_v2f esr_ld2(const float *p, i64 r, i64 n) {
  if (r + 1 < n) return *reinterpret_cast<const esr_v2f *>(p + r);
  esr_v2f v; v.x = p[r]; v.y = v.x; return v;
}
__device__ __forceinline__ void esr_st2(double *p, i64 r, i64 n, double x, double y) {
  if (r + 1 < n) { esr_v2d v; v.x = x; v.y = y; *reinterpret_cast<esr_v2d *>(p + r) = v; }
  else p[r] = x;
}
__device__ __forceinline__ void esr_st2(float *p, i64 r, i64 n, float x, float y) {
  if (r + 1 < n) { esr_v2f v; v.x = x; v.y = y; *reinterpret_cast<esr_v2f *>(p + r) = v; }
  else p[r] = x;
}
__device__ __forceinline__ void esr_st2nt(double *p, i64 r, i64 n, double x, double y) {
  if (r + 1 < n) {
    esr_v2d v; v.x = x; v.y = y;
    __builtin_nontemporal_store(v, reinterpret_cast<esr_v2d *>(p + r));
  } else __builtin_nontemporal_store(x, p + r);
}
__device__ __forceinline__ void esr_st2nt(float *p, i64 r, i64 n, float x, float y) {
  if (r + 1 < n) {
    esr_v2f v; v.x = x; v.y = y;
    __builtin_nontemporal_store(v, reinterpret_cast<esr_v2f *>(p + r));
  } else __builtin_nontemporal_store(x, p + r);
}
// per-row records of padded (ELL) edge arrays
typedef unsigned short esr_w4 __attribute__((ext_vector_type(4)));
typedef unsigned short esr_w8 __attribute__((ext_vector_type(8)));
typedef short esr_s4 __attribute__((ext_vector_type(4)));
typedef short esr_s8 __attribute__((ext_vector_type(8)));
typedef unsigned char esr_u4 __attribute__((ext_vector_type(4)));
typedef unsigned char esr_u8 __attribute__((ext_vector_type(8)));
typedef int esr_i4 __attribute__((ext_vector_type(4)));
typedef int esr_i8 __attribute__((ext_vector_type(8)));
typedef double esr_d4 __attribute__((ext_vector_type(4)));
typedef double esr_d8 __attribute__((ext_vector_type(8)));
typedef float esr_f4 __attribute__((ext_vector_type(4)));
typedef float esr_f8 __attribute__((ext_vector_type(8)));
// correctly rounded sqrt: device sqrt, then the Tuckerman rounding test
__device__ __forceinline__ double esr_sqrt(double x) {
  const double y = sqrt(x);
  if (!(x > 0x1p-900) || x > 0x1p+1000) return y;
  const long long yb = __builtin_bit_cast(long long, y);
  const double yu = __builtin_bit_cast(double, yb + 1);
  const double yd = __builtin_bit_cast(double, yb - 1);
  if (__builtin_fma(-y, yu, x) > 0.0) return yu;
  if (__builtin_fma(-y, yd, x) <= 0.0) return yd;
  return y;
}
__device__ __forceinline__ float esr_sqrt(float x) {
  const float y = sqrtf(x);
  if (!(x > 0x1p-100f) || x > 0x1p+100f) return y;
  const int yb = __builtin_bit_cast(int, y);
  const float yu = __builtin_bit_cast(float, yb + 1);
  const float yd = __builtin_bit_cast(float, yb - 1);
  if (__builtin_fmaf(-y, yu, x) > 0.0f) return yu;
  if (__builtin_fmaf(-y, yd, x) <= 0.0f) return yd;
  return y;
}
'''


class _Restage(Exception):
    """The staged edge arrays of a kernel exceed the LDS budget."""


class _RowsPerThread(Exception):
    """An array of a two-rows-per-thread kernel is not vector-aligned."""


class Kernel:
    def __init__(self, space, stage, name):
        self.space, self.stage, self.name = space, stage, name
        self.params: List[tuple] = []     # (cname, ctype, tensor)
        self.pindex: Dict[int, str] = {}
        self.access: Dict[str, set] = {}  # cname -> {'r', 'w'}
        self.n = 0
        self.src = ''
        self.staged = False
        self.lds_bytes = 0
        self.block = int(OPTIONS['block'])
        self.rpt = 1                      # rows per thread
        self.persist = 0                  # > 0: workgroups loop over row tiles
        self.luts: Dict[str, tuple] = {}  # dictionary param -> (LDS name, size)
        self.consts: Dict[str, str] = {}  # one-value dictionary param -> name

    def algorithmic_bytes(self) -> int:
        """Compulsory HBM bytes of one launch: every array the kernel reads
        counted once, every array it writes counted once (read-modify-write
        arrays twice); gathered source rows are counted once, as a read of
        the whole source array."""
        tot = 0
        for name, _, t in self.params:
            tot += t.numel() * t.element_size() * len(self.access.get(name, ()))
        return tot


class BoundTensors:
    """The tensors whose device addresses packed kernel arguments or a
    captured HIP graph hold, with those addresses. check() raises before a
    launch if one of them has been given new storage (its .data reassigned),
    which the launch would otherwise never see."""

    def __init__(self, tensors):
        self.pairs = [(t, t.data_ptr()) for t in tensors if t is not None]

    def check(self):
        for t, ptr in self.pairs:
            if t.data_ptr() != ptr:
                raise RuntimeError(
                    f'the storage of a compiled tensor {tuple(t.shape)} was replaced '
                    'after compile() (its .data reassigned); the compiled kernels '
                    'still address the old storage -- update it in place '
                    '(t.data.copy_(...)) or compile again')


class FusedEngine:
    """Executes the fused kernels of one forward; captured as a HIP graph
    after the first call."""

    def __init__(self, module, launches, keepalive, tensors=()):
        self.module = module
        self.launches = launches          # (fn, grid, argbytes, block)
        self.keepalive = keepalive
        self.graph = None
        self.calls = 0
        self.bound = BoundTensors(tensors)

    def _launch_all(self, stream):
        for fn, grid, buf, block in self.launches:
            nat.call('esr_launch', fn, grid, 1, 1, block, 1, 1, 0, stream,
                     ctypes.addressof(buf), ctypes.sizeof(buf))

    def forward(self):
        self.bound.check()
        self.calls += 1
        if self.graph is not None:
            nat.call('esr_graph_launch', self.graph.exec, nat.stream_ptr())
            return None
        self._launch_all(nat.stream_ptr())
        if self.calls >= 2:
            from ..numeric.solver.fused import _Graph
            self.graph = _Graph(self._launch_all)
        return None


class _DryBuf:
    """Stands in for an intermediate buffer while candidate plans are only
    costed (size and a unique, aligned address; never launched)."""
    _next = 1 << 40

    def __init__(self, n, itemsize):
        self.n, self.itemsize = n, itemsize
        _DryBuf._next += 1 << 12
        self.ptr = _DryBuf._next

    def data_ptr(self):
        return self.ptr

    def numel(self):
        return self.n

    def element_size(self):
        return self.itemsize


def _other_kernel_use(x: V, u: V) -> bool:
    """u reads x outside x's own kernel row loop, not through a Selector."""
    return u.kind != 'gather' and (u.stage != x.stage or u.space not in (x.space, None))


class _Lowering:
    def __init__(self, p: Program, prog: CompiledProgram, dry: bool = False):
        self.p, self.prog = p, prog
        self.ctype = 'double' if p.dtype == torch.float64 else 'float'
        self.device = prog.device
        self.kernels: List[Kernel] = []
        self.no_stage = False
        self.no_halo = False
        self.force_rpt1 = False
        self.dry = dry
        self.remat: set = set()
        self.enc: Dict[int, _Enc] = {}

    def _rowlocal(self, sel: _M.Selector, red: _M.Reducer) -> bool:
        """The Selector's index equals the Reducer's (it reads the row
        itself); computed once per pair, as every candidate plan asks."""
        a, b = getattr(sel, 'easier_host_idx', None), getattr(red, 'easier_host_idx', None)
        cache = sel.__dict__.setdefault('easier_rowlocal', {})
        # the entry keeps the objects it was computed from and is used only
        # while they are the same objects (ids alone could be reused)
        hit = cache.get(id(red))
        if hit is None or hit[0] is not red or hit[1] is not a or hit[2] is not b or \
                hit[3] != (red.easier_perm is None):
            same = a is not None and b is not None and a.shape == b.shape and \
                red.easier_perm is None and np.array_equal(a, b)
            hit = (red, a, b, red.easier_perm is None, same)
            cache[id(red)] = hit
        return hit[4]

    def _remat_set(self) -> set:
        """Ids of gathered elementwise values that their gathers recompute
        (OPTIONS['remat']). The value's expression is followed through
        elementwise nodes of its space down to constants, scalars, tensors
        (original contents) and values stored for another kernel anyway; at
        most two such arrays, and no tensor among them may be reassigned in
        or before the last kernel that gathers (its old contents are read at
        other rows)."""
        p = self.p
        out = set()
        for v in p.values:
            if v.kind != 'ew' or v.space is None or v.space in p.edge_spaces:
                continue
            gusers = [u for u in v.users if u.kind == 'gather']
            if not gusers or any(_other_kernel_use(v, u) for u in v.users):
                continue
            last = max(max(p.evals.get(u.id, {u.stage})) for u in gusers)
            arrays, tensors, seen, stack, ok = set(), set(), set(), [v], True
            while stack and ok:
                x = stack.pop()
                if x.id in seen:
                    continue
                seen.add(x.id)
                if x.kind in ('const', 'scalar'):
                    continue
                if x.kind == 'leaf' and x.space == v.space:
                    arrays.add(x.id)
                    tensors.add(x.obj)
                elif x is not v and x.space == v.space and x.kind in ('ew', 'reduce') and \
                        any(_other_kernel_use(x, u) for u in x.users):
                    arrays.add(x.id)
                elif x.kind == 'ew' and x.space in (v.space, None):
                    stack.extend(x.args)
                else:
                    ok = False
            if not ok or len(arrays) > 2:
                continue
            if any(st.obj in tensors and st.stage <= last for st in p.stores):
                continue
            out.add(v.id)
        return out

    def materialize(self):
        p = self.p
        self.remat = self._remat_set() if OPTIONS['remat'] else set()
        self.enc = _frozen_encoding(p, self.prog) if \
            getattr(self.prog, 'freeze', False) and OPTIONS['dict_max'] else {}
        for v in p.values:
            if v.kind not in ('ew', 'reduce', 'gather') or v.space is None or \
                    v.space in p.edge_spaces:
                continue
            need = any(_other_kernel_use(v, u) or
                       (u.kind == 'gather' and v.id not in self.remat) for u in v.users)
            v.mat = need
            v.buf, v.pack = None, None
            if need:
                n = self.prog.space_len[v.space]
                v.buf = _DryBuf(n, p.dtype.itemsize) if self.dry else \
                    torch.empty(n, dtype=p.dtype, device=self.device)
        if OPTIONS['pack_intermediates'] and OPTIONS['rows_per_thread'] == 1:
            # the values one kernel materialises share one row-interleaved
            # buffer: row r of value j at r * width + j
            groups: Dict[tuple, list] = {}
            for v in p.values:
                if v.mat and v.kind != 'leaf':
                    groups.setdefault((v.stage, v.space), []).append(v)
            for (stage, space), vs in groups.items():
                if len(vs) < 2:
                    continue
                n = self.prog.space_len[space]
                buf = _DryBuf(n * len(vs), p.dtype.itemsize) if self.dry else \
                    torch.empty(n * len(vs), dtype=p.dtype, device=self.device)
                for j, v in enumerate(vs):
                    v.buf, v.pack = buf, (len(vs), j)

    def build(self):
        p = self.p
        groups: Dict[tuple, list] = {}
        for v in p.values:
            if v.kind == 'store' or (v.mat and v.kind != 'leaf'):
                groups.setdefault((v.stage, v.space), []).append(v)
        srcs = []
        for k, (stage, space) in enumerate(sorted(groups, key=lambda x: (x[0], str(x[1])))):
            roots = groups[(stage, space)]
            kern = Kernel(space, stage, f'esr_fused_{k}')
            kern.n = self.prog.space_len[space]
            try:
                try:
                    self._emit_kernel(kern, roots)
                except _RowsPerThread:
                    self.force_rpt1 = True
                    kern = Kernel(space, stage, f'esr_fused_{k}')
                    kern.n = self.prog.space_len[space]
                    self._emit_kernel(kern, roots)
                    self.force_rpt1 = False
            except _NoHalo:
                self.no_halo = True
                kern = Kernel(space, stage, f'esr_fused_{k}')
                kern.n = self.prog.space_len[space]
                self._emit_kernel(kern, roots)
                self.no_halo = False
            except _Restage:
                # too many staged edge arrays for the LDS budget: this
                # kernel reads its edges straight from global memory
                self.no_stage = True
                kern = Kernel(space, stage, f'esr_fused_{k}')
                kern.n = self.prog.space_len[space]
                self._emit_kernel(kern, roots)
                self.no_stage = False
            self.kernels.append(kern)
            srcs.append(kern.src)
        return _CODE_HEADER + '\n'.join(srcs)

    # -- code emission --------------------------------------------------------
    def _param(self, kern, tensor, ctype, const=True):
        key = tensor.data_ptr()
        if key in kern.pindex:
            name = kern.pindex[key]
            if not const:       # read and written by this kernel
                k = int(name[3:])
                n_, _, t_ = kern.params[k]
                kern.params[k] = (n_, ctype + '*', t_)
        else:
            name = f'a.p{len(kern.params)}'
            kern.params.append((name, ('const ' if const else '') + ctype + '*', tensor))
            kern.pindex[key] = name
        kern.access.setdefault(name, set()).add('r' if const else 'w')
        return name

    def _emit_kernel(self, kern: Kernel, roots: List[V]):
        ct = self.ctype
        # two rows per thread (rows r0, r0 + 1): the row bodies are emitted
        # once per row, their row-indexed loads and stores shared as one
        # 2-element vector access; LDS edge staging keeps one row per thread
        rpt = 2 if (OPTIONS['rows_per_thread'] == 2 and not self.force_rpt1 and
                    not (OPTIONS['stage_edges'] and not self.no_stage)) else 1
        kern.rpt = rpt
        vt = 'esr_v2d' if ct == 'double' else 'esr_v2f'
        prologue: List[str] = []
        vloads: Dict[str, str] = {}

        class _Cx:
            rows: List[str]
            memo: Dict[int, str]
            groups: set
            R = 'r'
            comp = None
            sfx = ''
            gread = None
        cx = _Cx()

        def buf_of(v: V):
            if v.kind == 'leaf':
                return self._param(kern, v.obj.data, ct)
            if v.kind == 'scalar':
                return self._param(kern, v.obj.data, ct)
            assert v.mat and v.buf is not None
            return self._param(kern, v.buf, ct)

        def elem(v, row):
            """Index of v's element at `row` in its buffer."""
            if v is not None and v.pack is not None:
                return f'({row}) * {v.pack[0]} + {v.pack[1]}'
            return row

        def row_load(buf, v=None):
            """buf[row] in the current row context."""
            if cx.comp is None:
                return var(cx.rows, f'{buf}[{elem(v, cx.R)}]')
            if buf not in vloads:
                vloads[buf] = f'w{len(vloads)}'
                prologue.append(f'  const {vt} {vloads[buf]} = esr_ld2({buf}, r0, a.n);')
            return f'{vloads[buf]}.{cx.comp}'

        def var(stmts, expr, typ=None):
            name = f'v{len(memo_names)}'
            memo_names.append(name)
            stmts.append(f'  const {typ or ct} {name} = {expr};')
            return name
        memo_names: List[str] = []
        code_memo: Dict[tuple, str] = {}

        def leaf_at(t, pos, stmts):
            """Frozen tensor t at position `pos` through its dictionary (a
            kernel constant, or the LDS table at the row's code); None when t
            is read as an array."""
            enc = self.enc.get(id(t))
            if enc is None:
                return None
            lut = lut_name(enc)
            if enc.codes is None:
                return lut
            cp = self._param(kern, enc.codes, 'unsigned char')
            key = (cp, str(pos), id(stmts))
            if key not in code_memo:
                code_memo[key] = var(stmts, f'(int){cp}[{pos}]', 'int')
            return f'{lut}[{code_memo[key]}]'

        def lut_name(enc):
            """The kernel's name of a dictionary: its LDS table, or the
            constant of a one-value dictionary."""
            lut = self._param(kern, enc.lut, ct)
            if enc.codes is None:
                if lut not in kern.consts:
                    kern.consts[lut] = f'c{len(kern.consts)}'
                return kern.consts[lut]
            if not OPTIONS['dict_lds']:
                return lut
            if lut not in kern.luts:
                kern.luts[lut] = (f'L{len(kern.luts)}', enc.k)
            return kern.luts[lut][0]

        def ew_expr(v, a):
            op = v.op
            if op == 'add': return f'({a[0]} + {a[1]})'
            if op == 'sub': return f'({a[0]} - {a[1]})'
            if op == 'mul': return f'({a[0]} * {a[1]})'
            if op == 'div': return f'({a[0]} / {a[1]})'
            if op == 'neg': return f'(-{a[0]})'
            if op == 'bcast': return a[0]
            if op == 'abs': return f'fabs({a[0]})'
            if op == 'sqrt':
                return f'esr_sqrt({a[0]})'
            if op in ('exp', 'sin', 'cos', 'tanh', 'log'):
                return f'{op}({a[0]})'
            if op == 'sign':
                return (f'(({a[0]}) > ({ct})0 ? ({ct})1 : (({a[0]}) < ({ct})0 ? '
                        f'({ct})-1 : ({a[0]}) * ({ct})0))')
            if op == 'pow':
                e = v.args[1].const
                if e == 2: return f'({a[0]} * {a[0]})'
                if e == 3: return f'({a[0]} * {a[0]} * {a[0]})'
                if e == 0.5: return f'esr_sqrt({a[0]})'
                if e == -1: return f'(({ct})1 / {a[0]})'
                return f'pow({a[0]}, {a[1]})'
            if op in _PRED:
                sym = {'lt': '<', 'le': '<=', 'gt': '>', 'ge': '>=', 'eq': '==', 'ne': '!='}[op]
                return f'(({a[0]}) {sym} ({a[1]}))'
            if op == 'where': return f'(({a[0]}) ? ({a[1]}) : ({a[2]}))'
            if op == 'maximum':
                return f'(({a[0]} != {a[0]} || {a[0]} > {a[1]}) ? {a[0]} : {a[1]})'
            if op == 'minimum':
                return f'(({a[0]} != {a[0]} || {a[0]} < {a[1]}) ? {a[0]} : {a[1]})'
            raise Unsupported(op)

        def is_pred(v):
            return v.kind == 'ew' and v.op in _PRED

        def emit_at(y: V, gi: str, stmts: List[str], memo: Dict[tuple, str]) -> str:
            """y at row `gi` of its space, y a value its gathers recompute:
            its elementwise expression over the arrays it is made of, read
            at that row (same operations, same order)."""
            key = (y.id, gi)
            if key in memo:
                return memo[key]
            if y.kind == 'const':
                s = _fmt_const(y.const, ct)
            elif y.kind == 'scalar':
                s = emit_row(y)
            elif y.kind == 'leaf' and y.obj is not None and id(y.obj) in self.enc:
                if isinstance(gi, _LocalIdx):
                    s = var(stmts, cx.gread(('enc', y.id), lambda p, t=y.obj:
                                            leaf_at(t, p, cx.rows), gi, stmts))
                else:
                    s = var(stmts, leaf_at(y.obj, gi, stmts))
            elif y.kind == 'leaf' or y.mat:
                bn = buf_of(y)
                if isinstance(gi, _LocalIdx):
                    s = var(stmts, cx.gread(('buf', bn), lambda p, bn=bn, y=y:
                                            f'{bn}[{elem(y, p)}]', gi, stmts))
                else:
                    s = var(stmts, f'{bn}[{elem(y, gi)}]')
            elif y.kind == 'ew':
                a = [emit_at(x, gi, stmts, memo) for x in y.args]
                s = var(stmts, ew_expr(y, a), 'bool' if is_pred(y) else None)
            else:
                raise Unsupported(f'{y.kind} in a recomputed gather source')
            memo[key] = s
            return s

        def emit_row(v: V) -> str:
            if v.id in cx.memo:
                return cx.memo[v.id]
            if v.kind == 'const':
                s = _fmt_const(v.const, ct)
            elif v.kind == 'scalar':
                s = var(cx.rows, f'{buf_of(v)}[0]')
            elif v.kind == 'leaf' and cx.comp is None and id(v.obj) in self.enc:
                if v.space != kern.space:
                    raise Unsupported('row read from another space')
                s = var(cx.rows, leaf_at(v.obj, cx.R, cx.rows))
            elif v.kind == 'leaf' or (v.mat and v.stage < kern.stage) or \
                    (v.mat and v.space != kern.space):
                if v.space != kern.space:
                    raise Unsupported('row read from another space')
                s = row_load(buf_of(v), v)
            elif v.kind == 'gather':
                sel = v.obj
                src = v.args[0]
                idx = self._param(kern, sel.idx, 'int')
                if src.id in self.remat:
                    gi = var(cx.rows, f'{idx}[{cx.R}]', 'int')
                    s = emit_at(src, gi, cx.rows, {})
                else:
                    s = var(cx.rows, f'{buf_of(src)}[{elem(src, f"{idx}[{cx.R}]")}]')
            elif v.kind == 'reduce':
                emit_group(v)
                s = cx.memo[v.id]
            elif v.kind == 'ew':
                a = [emit_row(x) for x in v.args]
                s = var(cx.rows, ew_expr(v, a), 'bool' if is_pred(v) else None)
            else:
                raise Unsupported(v.kind)
            cx.memo[v.id] = s
            return s

        def emit_group(v: V):
            red = v.obj
            key = (id(red),)
            if key in cx.groups:
                return
            cx.groups.add(key)
            members = [u for u in self.p.values if u.kind == 'reduce' and u.obj is red
                       and u.stage == kern.stage and u.space == kern.space
                       and (u.id in needed)]
            hp = red.easier_host_rowptr
            counts = np.diff(hp)
            kmax = int(counts.max()) if counts.size else 0
            dense = 0 < kmax <= 8 and counts.mean() >= 0.5 * kmax
            # few non-empty rows (wall faces of a mesh): a bitmap of the
            # non-empty rows replaces the row pointers
            sparse = (not dense and OPTIONS['sparse_rows'] and counts.size >= 4096 and
                      np.count_nonzero(counts) * 8 <= counts.size)
            ell = (dense and red.easier_perm is None and kmax <= 8 and cx.comp is None and
                   getattr(self.prog, 'freeze', False) and OPTIONS['ell'] and
                   not (OPTIONS['stage_edges'] and not self.no_stage))
            S = 4 if kmax <= 4 else 8
            if sparse:
                bm, pre, crow = _sparse_rows(red)
                bmp = self._param(kern, bm, 'unsigned long long')
                prp = self._param(kern, pre, 'int')
                rp = self._param(kern, crow, 'int')
            elif ell:
                rp = None
                cntp = self._param(kern, _ell_counts(self.prog, red, self.device),
                                   'unsigned char')
            else:
                rp = self._param(kern, red.easier_rowptr, 'int')
            perm = None if red.easier_perm is None else \
                self._param(kern, red.easier_perm, 'int')
            body: List[str] = []
            emem: Dict[int, str] = {}
            amem: Dict[tuple, str] = {}
            gid = len(cx.groups)
            # edge arrays staged through LDS: param name -> (lds name, ctype)
            staged: Dict[str, tuple] = {}
            stage = dense and perm is None and OPTIONS['stage_edges'] and not self.no_stage
            nt = OPTIONS['nt_loads']
            B = kern.block

            ellv: Dict[str, str] = {}
            tags = {'short': 's', 'unsigned char': 'u', 'int': 'i', 'double': 'd',
                    'float': 'f', 'unsigned short': 'w'}
            # halo staging: key -> (LDS name, read(pos) -> C expression)
            hstage: Dict[tuple, tuple] = {}
            hinfo: Dict[int, tuple] = {}    # id(sel) -> (halo param, HS)
            use_halo = ell and OPTIONS['halo'] and not self.no_halo and \
                kern.block == _compiler.BLOCK_ROWS

            def gread(key, read, pos, stmts):
                """A value read at the gathered position `pos`: from the LDS
                stage when `pos` is a halo-local index (read(p) is the value's
                expression at global row p), else read(pos) itself."""
                if not (isinstance(pos, _LocalIdx)):
                    return read(pos)
                if key not in hstage:
                    hstage[key] = (f'H{gid}_{len(hstage)}', read, pos.sel)
                return f'{hstage[key][0]}[{pos}]'
            cx.gread = gread

            def ell_vec(t, ctype):
                """Row record of a padded edge array, loaded once per row
                (at row scope); slot k is <name>[k]."""
                name = self._param(kern, t, ctype)
                if name not in ellv:
                    ellv[name] = f'E{gid}_{len(ellv)}'
                    cx.rows.append(f'  const esr_{tags[ctype]}{S} {ellv[name]} = '
                                   f'*(const esr_{tags[ctype]}{S} *)({name} + (i64){cx.R} * {S});')
                return ellv[name]

            def ell_leaf(t) -> str:
                """Edge tensor t at slot k of the row."""
                enc = self.enc.get(id(t))
                if enc is not None and enc.codes is None:
                    return lut_name(enc)
                if enc is not None:
                    codes = _ell_array(self.prog, red, S, ('codes', id(enc.codes)),
                                       enc.codes.cpu().numpy(), self.device)
                    return f'{lut_name(enc)}[(int){ell_vec(codes, "unsigned char")}[k]]'
                raw = _ell_array(self.prog, red, S, ('raw', id(t)),
                                 t.data.detach().cpu().numpy(), self.device)
                return f'{ell_vec(raw, ct)}[k]'

            def edge_load(name, ctype):
                if stage:
                    if name not in staged:
                        staged[name] = (f's{gid}_{len(staged)}', ctype)
                    return f'{staged[name][0]}[jl]'
                if nt:
                    return f'__builtin_nontemporal_load(&{name}[e])'
                return f'{name}[e]'

            gidx: Dict[int, str] = {}

            def gather_index(sel, src):
                """The Selector's source row of edge e, loaded once per edge:
                int16 offsets from the edge's own row when they fit (dense
                rows in Reducer order), clamped into the source as the
                clamped slots of an empty row may belong to another row."""
                if id(sel) in gidx:
                    return gidx[id(sel)]
                hi = self.prog.space_len[src.space] - 1
                lay = _halo_layout(self.prog, sel, red, S, kern.block,
                                   min(int(OPTIONS['halo_max']), kern.block),
                                   self.device) if (use_halo and src.space == kern.space) \
                    else None
                if lay is not None:
                    loc, hal, hs = lay
                    hinfo[id(sel)] = (self._param(kern, hal, 'int'), hs)
                    li = var(body, f'(int){ell_vec(loc, "unsigned short")}[k]', 'int')
                    s = _LocalIdx(li, id(sel))
                    gidx[id(sel)] = s
                    return s
                if ell:
                    d = _delta16(sel, red) if OPTIONS['idx16'] else None
                    if d is not None:
                        o = _ell_array(self.prog, red, S, ('d16', id(sel)),
                                       d.cpu().numpy(), self.device)
                        s = var(body, f'esr_clampi((int){cx.R} + (int){ell_vec(o, "short")}[k], '
                                f'{hi})', 'int')
                    else:
                        ix = _ell_array(self.prog, red, S, ('idx', id(sel)),
                                        np.asarray(sel.easier_host_idx).astype(np.int32),
                                        self.device)
                        s = var(body, f'esr_clampi({ell_vec(ix, "int")}[k], {hi})', 'int')
                    gidx[id(sel)] = s
                    return s
                d = _delta16(sel, red) if (dense and perm is None and not stage and
                                           OPTIONS['idx16']) else None
                if d is not None:
                    dn = self._param(kern, d, 'short')
                    hi = self.prog.space_len[src.space] - 1
                    s = var(body, f'esr_clampi((int){cx.R} + (int){edge_load(dn, "short")}, {hi})',
                            'int')
                else:
                    idx = self._param(kern, sel.idx, 'int')
                    s = var(body, edge_load(idx, 'int'), 'int')
                gidx[id(sel)] = s
                return s

            def emit_edge(x: V) -> str:
                if x.id in emem:
                    return emem[x.id]
                if x.space != v.args[0].space:
                    s = emit_row(x)           # row value or constant, hoisted
                elif x.kind == 'leaf' and ell:
                    s = var(body, ell_leaf(x.obj))
                elif x.kind == 'leaf' and not stage and id(x.obj) in self.enc:
                    s = var(body, leaf_at(x.obj, 'e', body))
                elif x.kind == 'leaf':
                    s = var(body, edge_load(buf_of(x), ct))
                elif x.kind == 'gather':
                    src = x.args[0]
                    if self._rowlocal(x.obj, red):
                        s = emit_row(src) if src.kind != 'leaf' else \
                            emit_row_leaf(src)
                    elif src.id in self.remat:
                        s = emit_at(src, gather_index(x.obj, src), body, amem)
                    elif src.kind == 'leaf' and id(src.obj) in self.enc:
                        gi = gather_index(x.obj, src)
                        s = var(body, gread(('enc', src.id), lambda p, t=src.obj:
                                            leaf_at(t, p, cx.rows), gi, body))
                    else:
                        gi = gather_index(x.obj, src)
                        bn = buf_of(src)
                        s = var(body, gread(('buf', bn), lambda p, bn=bn, src=src:
                                            f'{bn}[{elem(src, p)}]', gi, body))
                elif x.kind == 'ew':
                    a = [emit_edge(y) for y in x.args]
                    s = var(body, ew_expr(x, a), 'bool' if is_pred(x) else None)
                else:
                    raise Unsupported(f'{x.kind} in an edge expression')
                emem[x.id] = s
                return s

            def emit_row_leaf(src):
                return emit_row(src)

            accs = []
            for u in members:
                accs.append(f'acc{u.id}{cx.sfx}')
            inner = [emit_edge(u.args[0]) for u in members]
            cx.rows.append('  ' + ' '.join(f'{ct} {a} = ({ct})0;' for a in accs))
            if stage:
                cap = B * kmax
                lds = sum(cap * (4 if c == 'int' else self.p.dtype.itemsize)
                          for c in [c for _, c in staged.values()])
                if kern.lds_bytes + lds > _LDS_BUDGET:
                    raise _Restage()
                kern.lds_bytes += lds
                kern.staged = True
                # the block's rows [rb, rb + B) own the contiguous edge
                # range [eb, eb + nl), nl <= B * kmax; copy it coalesced
                for name, (sn, c) in staged.items():
                    cx.rows.append(f'  __shared__ {c} {sn}[{cap}];')
                cx.rows.append(f'  const int eb{gid} = {rp}[rb], nl{gid} = '
                            f'{rp}[rb + {B} < a.n ? rb + {B} : a.n] - eb{gid};')
                ld = (lambda nm, i: f'__builtin_nontemporal_load(&{nm}[{i}])') if nt else \
                    (lambda nm, i: f'{nm}[{i}]')
                cx.rows.append(f'  for (int i = threadIdx.x; i < nl{gid}; i += {B}) {{')
                for name, (sn, c) in staged.items():
                    cx.rows.append(f'    {sn}[i] = {ld(name, f"eb{gid} + i")};')
                cx.rows.append('  }')
                # an empty range still gives every slot a valid (row 0) entry
                cx.rows.append(f'  if (nl{gid} == 0 && threadIdx.x == 0) {{ ' +
                            ' '.join(f'{sn}[0] = 0;' for sn, _ in staged.values()) + ' }')
                cx.rows.append('  __syncthreads();')
            if ell:
                for u, a in zip(members, accs):
                    body.append(f'  {a} = (k < cn{gid}) ? {a} + {inner[members.index(u)]} : {a};')
                if hstage:
                    # every thread of the tile stages its row, the first HS
                    # threads the tile's halo rows; then the edge slots read LDS
                    B = kern.block
                    isz = self.p.dtype.itemsize
                    lds = sum((B + hinfo[sid][1]) * isz for _, _, sid in hstage.values())
                    if kern.lds_bytes + lds > 64 * 1024:
                        raise _NoHalo()
                    kern.lds_bytes += lds
                    kern.staged = True
                    hrow = {}
                    for sid in dict.fromkeys(v[2] for v in hstage.values()):
                        hp, hs = hinfo[sid]
                        hrow[sid] = f'hr{gid}_{len(hrow)}'
                        cx.rows.append(f'  const int {hrow[sid]} = {hp}[(rb / {B}) * {hs} + '
                                       f'(threadIdx.x < {hs} ? threadIdx.x : 0)];')
                    for nm, read, sid in hstage.values():
                        hs = hinfo[sid][1]
                        cx.rows.append(f'  __shared__ {ct} {nm}[{B + hs}];')
                        own = read(cx.R)
                        cx.rows.append(f'  {nm}[threadIdx.x] = {own};')
                        far = read(hrow[sid])
                        cx.rows.append(f'  if (threadIdx.x < {hs}) {nm}[{B} + threadIdx.x] = {far};')
                    cx.rows.append('  __syncthreads();')
                cx.rows.append(f'  {{ const int cn{gid} = (int){cntp}[{cx.R}];')
                cx.rows.append(f'  #pragma unroll')
                cx.rows.append(f'  for (int k = 0; k < {kmax}; ++k) {{')
                cx.rows.extend('  ' + b for b in body)
                cx.rows.append('  } }')
            elif dense:
                # every row has at most kmax entries: evaluate kmax slots with
                # clamped (always valid) edge positions so all their loads
                # issue together, and fold slot k into the sums only when it
                # exists -- the same left-to-right order as the loop
                for u, a in zip(members, accs):
                    body.append(f'  {a} = (j < je) ? {a} + {inner[members.index(u)]} : {a};')
                cx.rows.append(f'  {{ const int jb = {rp}[{cx.R}], je = {rp}[{cx.R} + 1];')
                cx.rows.append(f'  #pragma unroll')
                cx.rows.append(f'  for (int k = 0; k < {kmax}; ++k) {{')
                cx.rows.append('    const int j = jb + k;')
                if stage:
                    cx.rows.append(f'    int jl = j - eb{gid}; jl = jl < nl{gid} ? jl : nl{gid} - 1; '
                                'jl = jl > 0 ? jl : 0;')
                else:
                    cx.rows.append('    const int jc = j < je ? j : (je > 0 ? je - 1 : 0);')
                    cx.rows.append(f'    const int e = {perm + "[jc]" if perm else "jc"};')
                cx.rows.extend('  ' + b for b in body)
                cx.rows.append('  } }')
            elif sparse:
                # compact row c of the row: the set bits of the bitmap before it
                for u, a in zip(members, accs):
                    body.append(f'  {a} = {a} + {inner[members.index(u)]};')
                R = cx.R
                cx.rows.append(f'  {{ const unsigned long long bw{gid} = {bmp}[{R} >> 6];')
                cx.rows.append(f'  if ((bw{gid} >> ({R} & 63)) & 1ull) {{')
                cx.rows.append(f'    const int c = {prp}[{R} >> 6] + '
                               f'__popcll(bw{gid} & ((1ull << ({R} & 63)) - 1ull));')
                cx.rows.append(f'    for (int j = {rp}[c]; j < {rp}[c + 1]; ++j) {{')
                cx.rows.append(f'      const int e = {perm + "[j]" if perm else "j"};')
                cx.rows.extend('    ' + b for b in body)
                cx.rows.append('    } } }')
            else:
                for u, a in zip(members, accs):
                    body.append(f'  {a} = {a} + {inner[members.index(u)]};')
                cx.rows.append(f'  for (int j = {rp}[{cx.R}]; j < {rp}[{cx.R} + 1]; ++j) {{')
                cx.rows.append(f'    const int e = {perm + "[j]" if perm else "j"};')
                cx.rows.extend('  ' + b for b in body)
                cx.rows.append('  }')
            for u, a in zip(members, accs):
                cx.memo[u.id] = a

        # values this kernel must compute (stop at values other kernels write)
        needed = set()

        def mark(v: V):
            if v.id in needed:
                return
            if v.kind in ('leaf', 'scalar', 'const'):
                return
            if v.mat and v.stage != kern.stage and v.kind != 'store':
                return
            needed.add(v.id)
            for a in v.args:
                mark(a)
        for rt in roots:
            needed.add(rt.id)
            for a in rt.args:
                mark(a)

        contexts = [('r', None, '')] if rpt == 1 else [('ra', 'x', 'a'), ('rb1', 'y', 'b')]
        ctx_stores = []
        body_rows: List[str] = []
        for R, comp, sfx in contexts:
            cx.rows, cx.memo, cx.groups = [], {}, set()
            cx.R, cx.comp, cx.sfx = R, comp, sfx
            stores = []
            for rt in roots:
                if rt.kind == 'store':
                    val = rt.args[0]
                    if val.mat and val.stage != kern.stage:
                        s = row_load(buf_of(val), val)
                    else:
                        s = emit_row(val)
                    stores.append((self._param(kern, rt.obj.data, ct, const=False), s,
                                   None))
                else:
                    s = emit_row(rt)
                    stores.append((self._param(kern, rt.buf, ct, const=False), s, rt))
            ctx_stores.append(stores)
            body_rows.extend(cx.rows)
        rows = prologue + body_rows
        nt_st = OPTIONS['nt_stores']
        if rpt == 1:
            if nt_st:
                st = [f'  __builtin_nontemporal_store({s}, &{d}[{elem(v, "r")}]);'
                      for d, s, v in ctx_stores[0]]
            else:
                st = [f'  {d}[{elem(v, "r")}] = {s};' for d, s, v in ctx_stores[0]]
        else:
            fn = 'esr_st2nt' if nt_st else 'esr_st2'
            st = [f'  {fn}({d}, r0, a.n, {sa}, {sb});'
                  for (d, sa, _), (_, sb, _) in zip(ctx_stores[0], ctx_stores[1])]
            # the paired accesses are 2-element vectors: their arrays must be
            # aligned to the vector (else this kernel keeps one row per thread)
            vec = 2 * self.p.dtype.itemsize
            names = set(vloads) | {d for d, _, _ in ctx_stores[0]}
            for cname, _, t in kern.params:
                if cname in names and t.data_ptr() % vec:
                    raise _RowsPerThread()
        params = ''.join(f'  {ctyp} __restrict__ p{i};\n'
                         for i, (_, ctyp, _) in enumerate(kern.params))
        if OPTIONS['xcd_group']:
            tile = f'esr_group_remap(blockIdx.x, gridDim.x, {int(OPTIONS["xcd_group"])})'
        elif OPTIONS['xcd_remap']:
            tile = 'esr_remap(blockIdx.x, gridDim.x)'
        else:
            tile = '(i64)blockIdx.x'
        if rpt == 2:
            head = (f'  const i64 r0 = ({tile} * {kern.block} + threadIdx.x) * 2;\n'
                    '  if (r0 >= a.n) return;\n'
                    '  const i64 ra = r0, rb1 = r0 + 1 < a.n ? r0 + 1 : r0;\n')
            tail = '\n'.join(st)
        elif kern.staged:
            # every thread of the block takes part in the LDS staging: rows
            # past the end recompute the last row and do not store
            head = (f'  const i64 rb = {tile} * {kern.block};\n'
                    '  const bool live = rb + threadIdx.x < a.n;\n'
                    '  const i64 r = live ? rb + threadIdx.x : a.n - 1;\n')
            tail = '  if (!live) return;\n' + '\n'.join(st)
        elif OPTIONS['persist'] and kern.luts:
            # at most `persist` workgroups loop over the row tiles: the
            # dictionary tables are staged once per workgroup
            kern.persist = int(OPTIONS['persist'])
            nt_ = f'((a.n + {kern.block} - 1) / {kern.block})'
            tl = tile.replace('blockIdx.x', 't_').replace('gridDim.x', nt_)
            head = (f'  for (i64 t_ = blockIdx.x; t_ < {nt_}; t_ += gridDim.x) {{\n'
                    f'  const i64 r = {tl} * {kern.block} + threadIdx.x;\n'
                    '  if (r >= a.n) continue;\n')
            tail = '\n'.join(st) + '\n  }'
        else:
            head = (f'  const i64 r = {tile} * {kern.block} + threadIdx.x;\n'
                    '  if (r >= a.n) return;\n')
            tail = '\n'.join(st)
        # dictionary tables to LDS (every thread takes part, before any
        # thread leaves), one-value dictionaries as constants
        pre = []
        for lname, (ln, k) in kern.luts.items():
            pre.append(f'  __shared__ {ct} {ln}[{k}];')
            pre.append(f'  for (int i = threadIdx.x; i < {k}; i += {kern.block}) '
                       f'{ln}[i] = {lname}[i];')
        for lname, cn in kern.consts.items():
            pre.append(f'  const {ct} {cn} = {lname}[0];')
        if kern.luts:
            pre.append('  __syncthreads();')
        head = ''.join(x + '\n' for x in pre) + head
        wpe = (f' __attribute__((amdgpu_waves_per_eu({int(OPTIONS["waves_per_eu"])})))'
               if OPTIONS['waves_per_eu'] else '')
        kern.src = (
            f'struct Args_{kern.name} {{\n{params}  i64 n;\n}};\n'
            f'extern "C" __global__ __launch_bounds__({kern.block}){wpe} void {kern.name}'
            f'(const Args_{kern.name} a) {{\n' + head + '\n'.join(rows) + '\n' +
            tail + '\n}\n')

    def pack(self, kern: Kernel) -> bytes:
        vals = [t.data_ptr() for (_, _, t) in kern.params]
        return struct.pack(f'{len(vals)}Qq', *vals, kern.n)


_MODULE_CACHE: Dict[str, object] = {}


def rtc_options(gfx: str) -> str:
    """hipRTC options: same floating-point contract as the prebuilt kernels
    (no FMA contraction, no packed-FP32 selection)."""
    return (f'--offload-arch={gfx} -O3 -std=c++17 -ffp-contract=off '
            '-Xclang -target-feature -Xclang -packed-fp32-ops')


def _compile_source(src: str):
    key = hashlib.sha1(src.encode()).hexdigest()
    if key in _MODULE_CACHE:
        return _MODULE_CACHE[key]
    arch = ctypes.create_string_buffer(256)
    nat.call('esr_device_arch', arch, 256)
    gfx = arch.value.decode().split(':')[0]
    mod = ctypes.c_void_p()
    nat.call('esr_rtc_compile', src.encode(), b'esr_fused.hip',
             rtc_options(gfx).encode(), ctypes.byref(mod))
    _MODULE_CACHE[key] = mod
    return mod


def lower_module(m: _M.Module, g: fx.Graph, prog: CompiledProgram):
    """(program, lowering, source) of a forward; raises Unsupported."""
    p = Program(m, g, prog)
    p.run()
    late = _late_stages(p, prog) if OPTIONS['late_eval'] else {}
    p.schedule(late)
    low = _Lowering(p, prog)
    low.materialize()
    src = low.build()
    return p, low, src


def _plan_bytes(p: Program, prog: CompiledProgram, late: Dict[int, int]) -> int:
    """Algorithmic bytes of all kernels of the plan with evaluation stages
    `late` (buffers stand-ins, nothing allocated or compiled)."""
    p.schedule(late)
    low = _Lowering(p, prog, dry=True)
    low.materialize()
    low.build()
    return sum(k.algorithmic_bytes() for k in low.kernels)


def _late_stages(p: Program, prog: CompiledProgram) -> Dict[int, int]:
    """OPTIONS['late_eval']: in program order, each elementwise row value
    whose users are all elementwise values of its space in later kernels is
    moved to the first of those kernels when the plan's bytes drop."""
    late: Dict[int, int] = {}
    best = _plan_bytes(p, prog, late)
    for x in list(p.values):
        if x.kind != 'ew' or x.space is None or x.space in p.edge_spaces or not x.users:
            continue
        if any(u.kind != 'ew' or u.space != x.space for u in x.users):
            continue
        target = min(u.stage for u in x.users)
        if target <= x.stage:
            continue
        trial = dict(late)
        trial[x.id] = target
        try:
            b = _plan_bytes(p, prog, trial)
        except Unsupported:
            b = None
        if b is not None and b < best:
            late, best = trial, b
        p.schedule(late)
    return late


def fuse_module(m: _M.Module, g: fx.Graph, prog: CompiledProgram):
    p, low, src = lower_module(m, g, prog)
    mod = _compile_source(src)
    launches = []
    for k in low.kernels:
        fn = ctypes.c_void_p()
        nat.call('esr_rtc_function', mod, k.name.encode(), ctypes.byref(fn))
        raw = low.pack(k)
        buf = (ctypes.c_char * len(raw)).from_buffer_copy(raw)
        grid = max(1, (k.n + k.block * k.rpt - 1) // (k.block * k.rpt))
        if k.persist:
            grid = min(grid, k.persist)
        launches.append((fn, grid, buf, k.block))
    keep = [v.buf for v in p.values if v.buf is not None]
    tensors = list({id(v.obj): v.obj for v in p.values
                    if v.kind in ('leaf', 'scalar', 'store')}.values())
    eng = FusedEngine(m, launches, keep, tensors)
    eng.source = src
    eng.kernels = low.kernels
    return eng


def try_fuse_program(prog: CompiledProgram):
    out = {}
    for m, g in zip(prog.modules, prog.graphs):
        try:
            out[m] = fuse_module(m, g, prog)
            logger.debug(f'{m.easier_hint_name}: fused into '
                         f'{len(out[m].kernels)} kernels')
        except Unsupported as e:
            logger.debug(f'{m.easier_hint_name}: node-by-node engine ({e})')
    return out
