"""Reads an HDF5 file with h5py and prints one JSON object: dataset path ->
[shape, dtype, sha256 of the values in native byte order]. tests/test_h5.py
runs it (with an interpreter that has h5py, e.g. /opt/conda/bin/python3.9)
on files written by easier_amd.core.h5, so that the HDF5 library itself
checks them."""
import hashlib
import json
import sys

import h5py
import numpy as np


def main(path):
    out = {}

    def visit(name, obj):
        if isinstance(obj, h5py.Dataset):
            a = np.ascontiguousarray(obj[()])
            a = a.astype(a.dtype.newbyteorder('='))
            out[name] = [list(a.shape), a.dtype.str.lstrip('<>|='),
                         hashlib.sha256(a.tobytes()).hexdigest()]
    with h5py.File(path, 'r') as f:
        f.visititems(visit)
    print(json.dumps(out, sort_keys=True))


if __name__ == '__main__':
    main(sys.argv[1])
