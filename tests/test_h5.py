"""The package's HDF5 reader and writer (easier_amd/core/h5.py) against
files the HDF5 library itself wrote and read:

* fixtures written by h5py 3.3 / HDF5 1.10.6 (tests/golden/make_h5_fixtures.py):
  the mesh schema of tutorial/create_triangular_mesh.py:29-35, every integer
  and floating dtype in both byte orders, scalar, empty and 3-d shapes,
  nested groups, a group of 40 members (several symbol-table nodes), compact
  storage, a dataset never written, datasets added in a second session
  (Tensor.save's mode 'a'), the libver='latest' format; read value for value
  against the arrays they were written from;
* a MATLAB 7.3 file (HDF5 with a 512-byte user block) that scipy ships as
  test data, against the same variable read by scipy.io.loadmat from the
  MATLAB 7.1 (non-HDF5) file of that test set;
* files this writer wrote, read back by h5py (and listed by h5dump) when an
  interpreter with h5py is installed (skipped otherwise).
"""
import hashlib
import json
import os
import shutil
import subprocess

import numpy as np
import pytest
import torch

from easier_amd.core import h5

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, 'golden')
H5PY_PYTHON = '/opt/conda/bin/python3.9'


def _expected():
    return np.load(os.path.join(GOLD, 'h5_expected.npz'))


def _native(a):
    return a.astype(a.dtype.newbyteorder('='))


@pytest.mark.parametrize('fname,prefix,count', [
    ('h5_mesh_tri4_n3.h5', 'mesh', 6), ('h5_types_default.h5', 'types', 59),
    ('h5_types_latest.h5', 'types', 8), ('h5_appended.h5', 'appended', 3)])
def test_reads_h5py_files(fname, prefix, count):
    exp = _expected()
    with h5.H5Reader(os.path.join(GOLD, fname)) as r:
        names = r.datasets()
        assert len(names) == count
        for n in names:
            want = exp[prefix + '__' + n.replace('/', '__')]
            got = r.read(n)
            assert got.dtype == _native(want).dtype and got.shape == want.shape, n
            assert np.array_equal(got, want), n


def test_reads_row_ranges():
    exp = _expected()
    with h5.H5Reader(os.path.join(GOLD, 'h5_types_default.h5')) as r:
        for n, lo, hi in (('f64', 5, 17), ('cube', 1, 3), ('i32', 0, 4), ('f64', 30, 99),
                          ('compact', 2, 5), ('unwritten', 1, 3), ('empty', 0, 0)):
            want = exp['types__' + n][lo:hi]
            got = r.read(n, lo, hi)
            assert got.shape == want.shape and np.array_equal(got, want), n


def test_mesh_schema_file_is_the_reference_mesh():
    """The h5py-written mesh file holds the oracle's tri4 (pinned against the
    reference's generator), and load_h5 returns the dict tri4 returns."""
    from oracle import mesh as omesh
    from easier_amd.workloads import mesh as wmesh
    m = wmesh.load_h5(os.path.join(GOLD, 'h5_mesh_tri4_n3.h5'))
    ref = omesh.tri4(3)
    for k in ('src', 'dst', 'cells', 'bcells', 'points', 'bpoints'):
        assert m[k].dtype == ref[k].dtype and np.array_equal(m[k], ref[k]), k


def test_refuses_what_it_does_not_implement():
    with h5.H5Reader(os.path.join(GOLD, 'h5_chunked.h5')) as r:
        with pytest.raises(h5.Unsupported, match='chunked'):
            r.read('chunked')
        assert np.array_equal(r.read('plain'), np.arange(4.0))
    with h5.H5Reader(os.path.join(GOLD, 'h5_dense.h5')) as r:
        with pytest.raises(h5.Unsupported, match='dense'):
            r.datasets()
    with pytest.raises(KeyError):
        h5.read(os.path.join(GOLD, 'h5_appended.h5'), 'nope')


def test_matlab_73_file_matches_scipy():
    scipy_io = pytest.importorskip('scipy.io')
    base = os.path.join(os.path.dirname(scipy_io.__file__), 'matlab', 'tests', 'data')
    p73 = os.path.join(base, 'testhdf5_7.4_GLNX86.mat')
    p71 = os.path.join(base, 'testdouble_7.1_GLNX86.mat')
    if not (os.path.exists(p73) and os.path.exists(p71)):
        pytest.skip('scipy test data not installed')
    with h5.H5Reader(p73) as r:
        assert r.datasets() == ['testdouble']
        got = r.read('testdouble')
    want = scipy_io.loadmat(p71)['testdouble']
    # MATLAB stores column-major: the HDF5 dataset is the transpose
    assert np.array_equal(got.T, want)
    assert np.allclose(want.ravel(), np.arange(9) * np.pi / 4)


def _arrays():
    rng = np.random.default_rng(5)
    a = {'src': np.arange(20, dtype=np.int64), 'points': rng.standard_normal((7, 2)),
         'f32': rng.standard_normal(5).astype(np.float32), 'i32': np.arange(-3, 3, dtype=np.int32),
         'scalar': np.array(2.5), 'empty': np.zeros((0, 3)), 'grp/a/b': np.ones((2, 3, 4)),
         'u8': np.arange(10, dtype=np.uint8), 'be': np.arange(4.0).astype('>f8'),
         'f16': np.arange(3).astype(np.float16), 'i16': np.arange(-4, 4, dtype=np.int16),
         'u64': np.array([0, 2 ** 64 - 1], np.uint64)}
    for k in range(70):                  # several symbol-table nodes
        a[f'many/x{k:02d}'] = np.full(2, k, np.int64)
    return a


def test_writer_round_trip(tmp_path):
    arrays = _arrays()
    path = str(tmp_path / 'w.h5')
    h5.write(path, arrays)
    raw = open(path, 'rb').read()
    assert raw[:8] == h5.SIGNATURE
    assert int.from_bytes(raw[40:48], 'little') == len(raw)        # end-of-file address
    with h5.H5Reader(path) as r:
        assert sorted(r.datasets()) == sorted(arrays)
        for k, a in arrays.items():
            got = r.read(k)
            assert got.dtype == _native(a).dtype and np.array_equal(got, a), k


def test_append_keeps_and_refuses_duplicates(tmp_path):
    path = str(tmp_path / 'a.h5')
    h5.append(path, {'x': np.arange(3.0)})
    h5.append(path, {'run/y': np.arange(4, dtype=np.int32)})
    with h5.H5Reader(path) as r:
        assert r.datasets() == ['run/y', 'x']
        assert np.array_equal(r.read('x'), np.arange(3.0))
    with pytest.raises(ValueError, match='exists'):
        h5.append(path, {'x': np.zeros(1)})
    # appending to a file h5py wrote keeps its datasets
    p2 = str(tmp_path / 'b.h5')
    shutil.copy(os.path.join(GOLD, 'h5_appended.h5'), p2)
    h5.append(p2, {'third': np.ones(2)})
    with h5.H5Reader(p2) as r:
        assert r.datasets() == ['first', 'run/step1', 'second', 'third']
        assert np.array_equal(r.read('second'), np.arange(6, dtype=np.float32).reshape(3, 2))


def test_hdf5_loader_reads_shape_dtype_rows():
    import easier_amd as esr
    path = os.path.join(GOLD, 'h5_mesh_tri4_n3.h5')
    ld = esr.hdf5(path, 'cells')
    assert ld.shape == (36, 3) and ld.dtype == torch.int64
    full = ld.load_cpu()
    assert torch.equal(ld.load_rows(4, 9), full[4:9])
    assert esr.hdf5(path, 'points', dtype=torch.float32).load_cpu().dtype == torch.float32


@pytest.mark.skipif(not os.path.exists(H5PY_PYTHON), reason='no interpreter with h5py')
def test_writer_output_read_by_h5py(tmp_path):
    arrays = _arrays()
    path = str(tmp_path / 'w.h5')
    h5.write(path, arrays)
    proc = subprocess.run([H5PY_PYTHON, os.path.join(GOLD, 'check_h5_with_h5py.py'), path],
                          capture_output=True, text=True, cwd=str(tmp_path))
    if proc.returncode != 0 and 'No module named' in proc.stderr:
        pytest.skip('the interpreter has no h5py')
    assert proc.returncode == 0, proc.stderr
    got = json.loads(proc.stdout)
    assert set(got) == set(arrays)
    for k, a in arrays.items():
        a = np.ascontiguousarray(_native(a))
        assert got[k] == [list(a.shape), a.dtype.str.lstrip('<>|='),
                          hashlib.sha256(a.tobytes()).hexdigest()], k
    h5dump = os.path.join(os.path.dirname(H5PY_PYTHON), 'h5dump')
    if os.path.exists(h5dump):
        proc = subprocess.run([h5dump, '-H', path], capture_output=True, text=True)
        assert proc.returncode == 0 and 'DATASET "x00"' in proc.stdout, proc.stderr
