"""Writes the HDF5 fixtures of tests/test_h5.py with h5py, the library the
reference reads and writes its mesh and tensor files with
(tutorial/create_triangular_mesh.py:29-35, easier/core/module.py:614-707).

Run with an interpreter that has h5py (this image: /opt/conda/bin/python3.9,
h5py 3.3.0 on HDF5 1.10.6):

    /opt/conda/bin/python3.9 tests/golden/make_h5_fixtures.py

Outputs (tests/golden/): h5_mesh_tri4_n3.h5 (the mesh schema, written as
create_triangular_mesh.py writes it), h5_types_default.h5 (every dtype and
shape case in h5py's default format) and h5_types_latest.h5 (a subset with
libver='latest': superblock version 3, version-2 object headers, compact
link groups), h5_appended.h5 (datasets added in a second session, as
Tensor.save's mode 'a' does), h5_dense.h5 (a libver='latest' group of 12
links, stored densely) and h5_chunked.h5 (a chunked, compressed dataset),
both refused by the reader; h5_expected.npz holds every array, for the
comparison, names with '/' replaced by '__'.
"""
import os
import sys

import h5py
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import mesh as omesh  # noqa: E402


def arrays():
    rng = np.random.default_rng(2024)
    out = {
        'f64': rng.standard_normal((37, 3)),
        'f32': rng.standard_normal(101).astype(np.float32),
        'f16': rng.standard_normal(9).astype(np.float16),
        'i64': rng.integers(-2 ** 62, 2 ** 62, 50, dtype=np.int64),
        'i32': rng.integers(-2 ** 31, 2 ** 31 - 1, (4, 5), dtype=np.int32),
        'i16': rng.integers(-2 ** 15, 2 ** 15 - 1, 7, dtype=np.int16),
        'i8': rng.integers(-128, 127, 11, dtype=np.int8),
        'u8': rng.integers(0, 255, 13, dtype=np.uint8),
        'u32': rng.integers(0, 2 ** 32 - 1, 6, dtype=np.uint32),
        'u64': rng.integers(0, 2 ** 63, 6, dtype=np.uint64),
        'be_f64': rng.standard_normal(8).astype('>f8'),
        'be_i32': rng.integers(-1000, 1000, 8).astype('>i4'),
        'scalar': np.array(3.25),
        'empty': np.zeros((0, 4), np.float64),
        'cube': rng.standard_normal((3, 4, 5)),
        'grp/sub/deep': np.arange(10, dtype=np.int64),
        'grp/x': rng.standard_normal(4),
    }
    for k in range(40):          # more members than one symbol table node holds
        out[f'many/d{k:02d}'] = np.full(3, k, np.int32)
    return out


def main():
    expected = {}
    m = omesh.tri4(3)
    path = os.path.join(HERE, 'h5_mesh_tri4_n3.h5')
    with h5py.File(path, 'w') as f:
        for k in ('src', 'dst', 'cells', 'bcells', 'points', 'bpoints'):
            f.create_dataset(k, data=m[k])
            expected[f'mesh__{k}'] = m[k]
    data = arrays()
    # libver='latest' stores a group of up to 8 links as link messages in its
    # object header (compact); beyond that in a fractal heap (dense), which
    # the reader refuses: the 'latest' file keeps every group at 8 members
    latest = ('f64', 'be_f64', 'scalar', 'empty', 'cube', 'grp/sub/deep')
    for tag, kw in (('default', {}), ('latest', {'libver': 'latest'})):
        with h5py.File(os.path.join(HERE, f'h5_types_{tag}.h5'), 'w', **kw) as f:
            for k, a in data.items():
                if tag == 'latest' and k not in latest:
                    continue
                f.create_dataset(k, data=a)
            # compact storage (raw data inside the object header)
            sid = h5py.h5s.create_simple((6,))
            dcpl = h5py.h5p.create(h5py.h5p.DATASET_CREATE)
            dcpl.set_layout(h5py.h5d.COMPACT)
            did = h5py.h5d.create(f.id, b'compact', h5py.h5t.NATIVE_DOUBLE, sid, dcpl=dcpl)
            compact = np.linspace(0.0, 1.0, 6)
            did.write(h5py.h5s.ALL, h5py.h5s.ALL, compact)
            # created but never written: reads as the fill value (zeros)
            f.create_dataset('unwritten', shape=(5,), dtype=np.float64)
    for k, a in data.items():
        expected['types__' + k.replace('/', '__')] = a
    expected['types__compact'] = np.linspace(0.0, 1.0, 6)
    expected['types__unwritten'] = np.zeros(5)
    path = os.path.join(HERE, 'h5_appended.h5')
    with h5py.File(path, 'w') as f:
        f.create_dataset('first', data=np.arange(5, dtype=np.float64))
    with h5py.File(path, 'a') as f:      # Tensor.save's mode
        f.create_dataset('second', dtype=np.float32, shape=(3, 2))
        f['second'][...] = np.arange(6, dtype=np.float32).reshape(3, 2)
        f.create_dataset('run/step1', data=np.array([7, 8, 9], np.int64))
    expected['appended__first'] = np.arange(5, dtype=np.float64)
    expected['appended__second'] = np.arange(6, dtype=np.float32).reshape(3, 2)
    expected['appended__run__step1'] = np.array([7, 8, 9], np.int64)
    with h5py.File(os.path.join(HERE, 'h5_dense.h5'), 'w', libver='latest') as f:
        for k in range(12):
            f.create_dataset(f'd{k:02d}', data=np.arange(3.0))
    with h5py.File(os.path.join(HERE, 'h5_chunked.h5'), 'w') as f:
        f.create_dataset('chunked', data=np.arange(100.0), chunks=(10,), compression='gzip')
        f.create_dataset('plain', data=np.arange(4.0))
    np.savez(os.path.join(HERE, 'h5_expected.npz'), **expected)
    print('wrote fixtures to', HERE)


if __name__ == '__main__':
    main()
