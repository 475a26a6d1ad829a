"""HDF5 files without h5py: a reader and a writer for the part of the format
that mesh and tensor files use.

The reference reads meshes and tensors with h5py (esr.hdf5 and H5DataLoader,
easier/core/module.py:35-52, easier/core/runtime/data_loader.py:418-751;
the mesh schema of tutorial/create_triangular_mesh.py:29-35) and writes
tensors with Tensor.save (module.py:614-707). h5py is not installed here, so
this module implements that part of the HDF5 file format (The HDF Group,
"HDF5 File Format Specification Version 3.0") directly:

Reading: superblock versions 0-3 (also behind a user block), object headers
version 1 and 2 with continuation blocks, groups stored as symbol tables
(version-1 B-tree + local heap, what h5py writes by default) or as compact
link messages (what libver='latest' writes for small groups); datasets with
a simple or scalar dataspace, fixed-point (signed/unsigned, 1-8 bytes) or
IEEE floating-point (2, 4, 8 bytes) elements in either byte order, and
contiguous or compact storage, read whole or as a range of leading rows
(without loading the rest of the file). Chunked (possibly compressed)
datasets and dense (fractal-heap) groups raise Unsupported.

Writing: a new file in the oldest format (superblock version 0, version-1
object headers, symbol-table groups), one contiguous dataset per array,
nested groups from '/'-separated paths; append() adds datasets to a file
this module can read by rewriting it.
"""
import os
from typing import Dict, Iterator, List, Optional, Tuple

import numpy as np

SIGNATURE = b'\x89HDF\r\n\x1a\n'


class Unsupported(ValueError):
    """A valid HDF5 feature outside the part this module implements."""


class FormatError(ValueError):
    """Bytes that are not the HDF5 structure expected at that place."""


# message types (format specification IV.A.2)
_NIL, _DATASPACE, _LINK_INFO, _DATATYPE, _FILL_OLD, _FILL, _LINK = 0, 1, 2, 3, 4, 5, 6
_LAYOUT, _FILTERS, _CONTINUATION, _SYMBOL_TABLE = 8, 11, 16, 17


class Dataset:
    """A dataset's metadata: shape, numpy dtype and where its bytes are."""

    def __init__(self, name, shape, dtype, address, nbytes, compact=None):
        self.name, self.shape, self.dtype = name, tuple(shape), dtype
        self.address, self.nbytes, self.compact = address, nbytes, compact

    def __repr__(self):
        return f'Dataset({self.name!r}, shape={self.shape}, dtype={self.dtype})'


class H5Reader:
    """An HDF5 file opened for reading (use as a context manager)."""

    def __init__(self, path: str):
        self.path = os.path.expanduser(path)
        self.f = open(self.path, 'rb')
        try:
            self._superblock()
        except Exception:
            self.f.close()
            raise

    def close(self):
        self.f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- low level ------------------------------------------------------------
    def _raw(self, addr: int, n: int) -> bytes:
        self.f.seek(self.base + addr)
        b = self.f.read(n)
        if len(b) != n:
            raise FormatError(f'{self.path}: {n} bytes at {addr} run past the end of the file')
        return b

    def _uint(self, b: bytes, off: int, size: int) -> int:
        return int.from_bytes(b[off:off + size], 'little')

    def _undef(self, v: int, size: int) -> bool:
        return v == (1 << (8 * size)) - 1

    def _superblock(self):
        # the superblock sits at 0 or, behind a user block, at 512, 1024, ...
        size = os.path.getsize(self.path)
        at = 0
        while True:
            self.f.seek(at)
            if self.f.read(8) == SIGNATURE:
                break
            at = 512 if at == 0 else 2 * at
            if at + 8 > size:
                raise FormatError(f'{self.path}: no HDF5 signature')
        self.f.seek(at)
        b = self.f.read(256)
        ver = b[8]
        if ver in (0, 1):
            self.O, self.L = b[13], b[14]
            p = 24 + (4 if ver == 1 else 0)
            O = self.O
            base = self._uint(b, p, O)
            root_entry = p + 4 * O
            self.root = self._uint(b, root_entry + O, O)
        elif ver in (2, 3):
            self.O, self.L = b[9], b[10]
            O = self.O
            base = self._uint(b, 12, O)
            self.root = self._uint(b, 12 + 3 * O, O)
        else:
            raise Unsupported(f'{self.path}: superblock version {ver}')
        if self.O not in (2, 4, 8) or self.L not in (2, 4, 8):
            raise FormatError(f'{self.path}: offset/length sizes {self.O}/{self.L}')
        # addresses are relative to the base address (the superblock's own
        # position when a user block precedes it)
        self.base = base if base else at

    # -- object headers -------------------------------------------------------
    def _messages(self, addr: int) -> List[Tuple[int, bytes]]:
        head = self._raw(addr, 4)
        if head == b'OHDR':
            return self._messages_v2(addr)
        return self._messages_v1(addr)

    def _messages_v1(self, addr: int) -> List[Tuple[int, bytes]]:
        b = self._raw(addr, 16)
        if b[0] != 1:
            raise FormatError(f'object header at {addr}: version {b[0]}')
        nmsg = self._uint(b, 2, 2)
        size = self._uint(b, 8, 4)
        blocks = [(addr + 16, size)]
        out = []
        while blocks and len(out) < nmsg:
            start, length = blocks.pop(0)
            data = self._raw(start, length)
            p = 0
            while p + 8 <= length and len(out) < nmsg:
                mtype, msize = self._uint(data, p, 2), self._uint(data, p + 2, 2)
                body = data[p + 8:p + 8 + msize]
                if mtype == _CONTINUATION:
                    blocks.append((self._uint(body, 0, self.O), self._uint(body, self.O, self.L)))
                out.append((mtype, body))
                p += 8 + ((msize + 7) & ~7)
        return out

    def _messages_v2(self, addr: int) -> List[Tuple[int, bytes]]:
        b = self._raw(addr, 64)
        if b[4] != 2:
            raise FormatError(f'object header at {addr}: version {b[4]}')
        flags = b[5]
        p = 6 + (16 if flags & 0x20 else 0) + (4 if flags & 0x10 else 0)
        csize = 1 << (flags & 3)
        chunk0 = self._uint(b, p, csize)
        p += csize
        blocks = [(addr + p, chunk0)]
        order = bool(flags & 0x04)
        out = []
        while blocks:
            start, length = blocks.pop(0)
            data = self._raw(start, length)
            q = 0
            hdr = 6 if order else 4
            while q + hdr <= length:
                mtype, msize = data[q], self._uint(data, q + 1, 2)
                body = data[q + hdr:q + hdr + msize]
                if mtype == _CONTINUATION:
                    caddr = self._uint(body, 0, self.O)
                    clen = self._uint(body, self.O, self.L)
                    if self._raw(caddr, 4) != b'OCHK':
                        raise FormatError(f'continuation block at {caddr}: no OCHK')
                    blocks.append((caddr + 4, clen - 8))       # signature, checksum
                out.append((mtype, body))
                q += hdr + msize
        return out

    # -- groups ---------------------------------------------------------------
    def _links(self, addr: int) -> Dict[str, int]:
        """Name -> object header address of the members of the group whose
        object header is at addr."""
        msgs = self._messages(addr)
        links = {}
        for mtype, body in msgs:
            if mtype == _SYMBOL_TABLE:
                btree = self._uint(body, 0, self.O)
                heap = self._uint(body, self.O, self.O)
                links.update(self._symbol_table(btree, heap))
            elif mtype == _LINK:
                name, target = self._link(body)
                if target is not None:
                    links[name] = target
            elif mtype == _LINK_INFO:
                heap = self._uint(body, 2 + (8 if body[1] & 1 else 0), self.O)
                if not self._undef(heap, self.O):
                    raise Unsupported('groups with dense link storage (fractal heap)')
        return links

    def _link(self, b: bytes):
        flags = b[1]
        p = 2
        ltype = 0
        if flags & 0x08:
            ltype = b[p]
            p += 1
        if flags & 0x04:
            p += 8
        if flags & 0x10:
            p += 1
        nsize = 1 << (flags & 3)
        nlen = self._uint(b, p, nsize)
        p += nsize
        name = b[p:p + nlen].decode('utf-8')
        p += nlen
        if ltype != 0:           # soft and external links are not followed
            return name, None
        return name, self._uint(b, p, self.O)

    def _heap_data(self, heap: int) -> Tuple[int, int]:
        b = self._raw(heap, 8 + 2 * self.L + self.O)
        if b[:4] != b'HEAP':
            raise FormatError(f'local heap at {heap}: no HEAP signature')
        size = self._uint(b, 8, self.L)
        return self._uint(b, 8 + 2 * self.L, self.O), size

    def _symbol_table(self, btree: int, heap: int) -> Dict[str, int]:
        data_addr, data_size = self._heap_data(heap)
        names = self._raw(data_addr, data_size)
        out = {}
        for snod in self._btree_children(btree):
            b = self._raw(snod, 8)
            if b[:4] != b'SNOD':
                raise FormatError(f'symbol table node at {snod}: no SNOD signature')
            n = self._uint(b, 6, 2)
            esize = 2 * self.O + 24
            ent = self._raw(snod + 8, n * esize)
            for k in range(n):
                off = self._uint(ent, k * esize, self.O)
                obj = self._uint(ent, k * esize + self.O, self.O)
                end = names.index(b'\0', off)
                out[names[off:end].decode('utf-8')] = obj
        return out

    def _btree_children(self, addr: int) -> Iterator[int]:
        b = self._raw(addr, 8 + 2 * self.O)
        if b[:4] != b'TREE':
            raise FormatError(f'B-tree node at {addr}: no TREE signature')
        if b[4] != 0:
            raise FormatError(f'B-tree node at {addr}: type {b[4]} is not a group node')
        level, used = b[5], self._uint(b, 6, 2)
        p = 8 + 2 * self.O
        body = self._raw(addr + p, used * (self.L + self.O) + self.L)
        for k in range(used):
            child = self._uint(body, k * (self.L + self.O) + self.L, self.O)
            if level > 0:
                yield from self._btree_children(child)
            else:
                yield child

    def _object(self, path: str) -> int:
        addr = self.root
        for part in [p for p in path.split('/') if p]:
            links = self._links(addr)
            if part not in links:
                raise KeyError(f'{self.path}: no object {path!r}')
            addr = links[part]
        return addr

    # -- datasets -------------------------------------------------------------
    def _dtype(self, b: bytes) -> np.dtype:
        cls, bits, size = b[0] & 0x0F, b[1:4], self._uint(b, 4, 4)
        if cls == 0:
            order = '>' if bits[0] & 1 else '<'
            if size not in (1, 2, 4, 8):
                raise Unsupported(f'integers of {size} bytes')
            return np.dtype(f'{order}{"i" if bits[0] & 0x08 else "u"}{size}')
        if cls == 1:
            if bits[0] & 0x40:
                raise Unsupported('VAX-order floating point')
            order = '>' if bits[0] & 1 else '<'
            if size not in (2, 4, 8):
                raise Unsupported(f'floating point of {size} bytes')
            return np.dtype(f'{order}f{size}')
        raise Unsupported(f'datatype class {cls} (only integers and floating point)')

    def _dataspace(self, b: bytes) -> Tuple[int, ...]:
        ver, nd = b[0], b[1]
        if ver == 1:
            p = 8
        elif ver == 2:
            if b[3] == 2:
                raise Unsupported('null dataspace')
            p = 4
        else:
            raise Unsupported(f'dataspace version {ver}')
        return tuple(self._uint(b, p + k * self.L, self.L) for k in range(nd))

    def _layout(self, b: bytes, shape, itemsize):
        ver = b[0]
        if ver in (3, 4):
            cls = b[1]
            if cls == 0:
                n = self._uint(b, 2, 2)
                return None, n, b[4:4 + n]
            if cls == 1:
                return self._uint(b, 2, self.O), self._uint(b, 2 + self.O, self.L), None
            raise Unsupported('chunked (or virtual) dataset storage; store the dataset '
                              'contiguous (h5py: create_dataset without chunks or filters)')
        if ver in (1, 2):
            nd, cls = b[1], b[2]
            p = 8
            if cls == 2:
                raise Unsupported('chunked dataset storage')
            addr = None
            if cls == 1:
                addr = self._uint(b, p, self.O)
                p += self.O
            p += 4 * nd
            if cls == 0:
                n = self._uint(b, p, 4)
                return None, n, b[p + 4:p + 4 + n]
            return addr, int(np.prod(shape, dtype=np.int64)) * itemsize, None
        raise Unsupported(f'data layout version {ver}')

    def dataset(self, path: str) -> Dataset:
        msgs = dict()
        for mtype, body in self._messages(self._object(path)):
            msgs.setdefault(mtype, body)
        if _DATASPACE not in msgs or _DATATYPE not in msgs or _LAYOUT not in msgs:
            raise KeyError(f'{self.path}: {path!r} is not a dataset')
        shape = self._dataspace(msgs[_DATASPACE])
        dtype = self._dtype(msgs[_DATATYPE])
        addr, nbytes, compact = self._layout(msgs[_LAYOUT], shape, dtype.itemsize)
        want = int(np.prod(shape, dtype=np.int64)) * dtype.itemsize
        if compact is None and addr is not None and self._undef(addr, self.O):
            addr = None                  # never written: the fill value (zeros)
        elif nbytes < want:
            raise FormatError(f'{path!r}: {nbytes} bytes stored for {want}')
        return Dataset(path, shape, dtype, addr, nbytes, compact)

    def read(self, path: str, start: int = 0, stop: Optional[int] = None) -> np.ndarray:
        """Rows [start, stop) of the dataset (all of it by default), in native
        byte order."""
        d = self.dataset(path)
        if not d.shape:
            rows, row_elems = 1, 1
            start, stop = 0, 1
        else:
            rows = d.shape[0]
            row_elems = int(np.prod(d.shape[1:], dtype=np.int64))
            stop = rows if stop is None else min(int(stop), rows)
            start = min(max(int(start), 0), stop)
        out_shape = d.shape if not d.shape else (stop - start,) + d.shape[1:]
        count = (stop - start) * row_elems
        rowbytes = row_elems * d.dtype.itemsize
        if d.compact is not None:
            raw = d.compact[start * rowbytes:stop * rowbytes]
        elif d.address is None:
            return np.zeros(out_shape, d.dtype.newbyteorder('='))
        else:
            raw = self._raw(d.address + start * rowbytes, count * d.dtype.itemsize)
        arr = np.frombuffer(raw, dtype=d.dtype, count=count).reshape(out_shape)
        return arr.astype(d.dtype.newbyteorder('='), copy=True)

    def datasets(self, group: str = '/') -> List[str]:
        """Paths of every dataset under group (depth first, sorted names)."""
        out = []

        def walk(addr, prefix):
            for name in sorted(self._links(addr)):
                child = self._links(addr)[name]
                path = f'{prefix}/{name}'
                kinds = {t for t, _ in self._messages(child)}
                if _LAYOUT in kinds:
                    out.append(path.lstrip('/'))
                elif kinds & {_SYMBOL_TABLE, _LINK, _LINK_INFO}:
                    walk(child, path)
        walk(self._object(group), '' if group.strip('/') == '' else '/' + group.strip('/'))
        return out


def read(path: str, dataset: str) -> np.ndarray:
    with H5Reader(path) as r:
        return r.read(dataset)


# ---------------------------------------------------------------------------
# writer
# ---------------------------------------------------------------------------
_O = 8                  # size of offsets and lengths written
_UNDEF = (1 << 64) - 1
_LEAF_K = 4             # group leaf node K: up to 8 entries per symbol table node
_INTERNAL_K = 16        # group internal node K: up to 32 children per B-tree node


def _u(v: int, n: int) -> bytes:
    return int(v).to_bytes(n, 'little')


def _pad8(b: bytes) -> bytes:
    return b + b'\0' * (-len(b) % 8)


def _datatype_message(dt: np.dtype) -> bytes:
    dt = np.dtype(dt)
    size = dt.itemsize
    if dt.kind in 'iu':
        bits = bytes([0x08 if dt.kind == 'i' else 0x00, 0, 0])
        return bytes([0x10]) + bits + _u(size, 4) + _u(0, 2) + _u(8 * size, 2)
    if dt.kind == 'f' and size in (2, 4, 8):
        exp_bits, man_bits, bias = {2: (5, 10, 15), 4: (8, 23, 127), 8: (11, 52, 1023)}[size]
        bits = bytes([0x20, 8 * size - 1, 0])     # little endian, implied MSB, sign bit
        return bytes([0x11]) + bits + _u(size, 4) + _u(0, 2) + _u(8 * size, 2) + \
            bytes([man_bits, exp_bits, 0, man_bits]) + _u(bias, 4)
    raise Unsupported(f'arrays of dtype {dt} (only integers and floating point are written)')


def _message(mtype: int, body: bytes) -> bytes:
    body = _pad8(body)
    return _u(mtype, 2) + _u(len(body), 2) + bytes([0, 0, 0, 0]) + body


def _object_header(messages: List[bytes]) -> bytes:
    data = b''.join(messages)
    return bytes([1, 0]) + _u(len(messages), 2) + _u(1, 4) + _u(len(data), 4) + \
        b'\0' * 4 + data


class _Group:
    def __init__(self):
        self.members: Dict[str, object] = {}


def _build_tree(arrays: Dict[str, np.ndarray]) -> _Group:
    root = _Group()
    for path, arr in arrays.items():
        parts = [p for p in path.split('/') if p]
        if not parts:
            raise ValueError(f'empty dataset path {path!r}')
        g = root
        for p in parts[:-1]:
            nxt = g.members.setdefault(p, _Group())
            if not isinstance(nxt, _Group):
                raise ValueError(f'{path!r}: {p!r} is a dataset, not a group')
            g = nxt
        if parts[-1] in g.members:
            raise ValueError(f'dataset {path!r} given twice')
        g.members[parts[-1]] = np.asarray(arr)
    return root


def write(path: str, arrays: Dict[str, np.ndarray]) -> None:
    """Write a new HDF5 file (replacing any file at path) holding each array
    as a contiguous dataset; '/' in a name makes nested groups."""
    root = _build_tree(arrays)
    out = bytearray(b'\0' * 96)          # superblock, filled in last

    def alloc(b: bytes) -> int:
        addr = len(out)
        out.extend(_pad8(b))
        return addr

    def dataset(arr: np.ndarray) -> int:
        a = np.asarray(arr)              # (ascontiguousarray would make a scalar 1-d)
        dt = a.dtype.newbyteorder('<')
        a = np.require(a.astype(dt, copy=False), requirements='C')
        data_addr = alloc(a.tobytes()) if a.nbytes else _UNDEF
        space = bytes([1, a.ndim, 0, 0]) + b'\0' * 4 + b''.join(_u(s, _O) for s in a.shape)
        # version 2, late allocation, write when set, defined with size 0
        # (the library's default fill: zeros)
        fill = bytes([2, 2, 2, 1]) + _u(0, 4)
        layout = bytes([3, 1]) + _u(data_addr, _O) + _u(a.nbytes, _O)
        return alloc(_object_header([_message(_DATASPACE, space),
                                     _message(_DATATYPE, _datatype_message(dt)),
                                     _message(_FILL, fill),
                                     _message(_LAYOUT, layout)]))

    def group(g: _Group) -> Tuple[int, int, int]:
        """(object header, B-tree, local heap) addresses of g."""
        names = sorted(g.members, key=lambda s: s.encode('utf-8'))
        entries = []
        for name in names:
            m = g.members[name]
            if isinstance(m, _Group):
                oh, bt, hp = group(m)
                entries.append((name, oh, 1, _u(bt, _O) + _u(hp, _O)))
            else:
                entries.append((name, dataset(m), 0, b'\0' * 16))
        # local heap: the empty string at offset 0, then every name
        heap = bytearray(b'\0' * 8)
        offsets = {}
        for name in names:
            offsets[name] = len(heap)
            heap.extend(_pad8(name.encode('utf-8') + b'\0'))
        heap_data = alloc(bytes(heap))
        # free-list offset 1: no free block (offsets of free blocks are
        # multiples of 8, so the library uses 1 for the empty list)
        heap_addr = alloc(b'HEAP' + bytes([0, 0, 0, 0]) + _u(len(heap), _O) + _u(1, _O) +
                          _u(heap_data, _O))
        # symbol table nodes of at most 2 K entries, sorted by name
        per = 2 * _LEAF_K
        chunks = [entries[k:k + per] for k in range(0, len(entries), per)] or [[]]
        if len(chunks) > 2 * _INTERNAL_K:
            raise Unsupported(f'more than {2 * _INTERNAL_K * per} members in one group')
        snods = []
        for ch in chunks:
            body = bytearray(b'SNOD' + bytes([1, 0]) + _u(len(ch), 2))
            for name, oh, cache, scratch in ch:
                body.extend(_u(offsets[name], _O) + _u(oh, _O) + _u(cache, 4) + b'\0' * 4 +
                            scratch)
            body.extend(b'\0' * ((per - len(ch)) * (2 * _O + 24)))
            snods.append((alloc(bytes(body)), offsets[ch[-1][0]] if ch else 0))
        tree = bytearray(b'TREE' + bytes([0, 0]) + _u(len(snods), 2) + _u(_UNDEF, _O) +
                         _u(_UNDEF, _O))
        tree.extend(_u(0, _O))                               # key 0: the empty name
        for addr, last in snods:
            tree.extend(_u(addr, _O) + _u(last, _O))         # child, then its last name
        tree.extend(b'\0' * ((2 * _INTERNAL_K - len(snods)) * 2 * _O))
        btree = alloc(bytes(tree))
        oh = alloc(_object_header([_message(_SYMBOL_TABLE, _u(btree, _O) + _u(heap_addr, _O))]))
        return oh, btree, heap_addr

    root_oh, root_bt, root_hp = group(root)
    sb = bytearray(SIGNATURE)
    sb += bytes([0, 0, 0, 0, 0, _O, _O, 0])
    sb += _u(_LEAF_K, 2) + _u(_INTERNAL_K, 2) + _u(0, 4)
    sb += _u(0, _O) + _u(_UNDEF, _O) + _u(len(out), _O) + _u(_UNDEF, _O)
    sb += _u(0, _O) + _u(root_oh, _O) + _u(1, 4) + b'\0' * 4 + _u(root_bt, _O) + _u(root_hp, _O)
    assert len(sb) == 96
    out[:96] = sb
    path = os.path.expanduser(path)
    tmp = path + '.tmp'
    with open(tmp, 'wb') as f:
        f.write(bytes(out))
    os.replace(tmp, path)


def append(path: str, arrays: Dict[str, np.ndarray]) -> None:
    """Add datasets to the file at path (created if absent), the reference's
    h5py mode 'a'. Existing datasets are kept (the file is rewritten); a name
    that exists already raises, as creating it again in h5py does."""
    path = os.path.expanduser(path)
    merged: Dict[str, np.ndarray] = {}
    if os.path.exists(path):
        with H5Reader(path) as r:
            for name in r.datasets():
                merged[name] = r.read(name)
    for name, arr in arrays.items():
        key = name.strip('/')
        if key in merged:
            raise ValueError(f'{path}: dataset {key!r} exists already')
        merged[key] = np.asarray(arr)
    write(path, merged)
