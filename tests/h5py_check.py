"""Cross-check helper for test_h5.py, run by a separate interpreter that has h5py.

  h5py_check.py dump FILE      -> JSON of every group / dataset / attribute h5py sees in FILE
  h5py_check.py make FILE LIBVER -> writes a Mosaic-style file with h5py (libver 'latest' or
                                   'earliest') holding the values of sample()
"""
import json
import sys

import h5py
import numpy as np


def sample():
    return {'data': np.arange(24, dtype=np.float32).reshape(2, 3, 4) / 7, 'name': 'vp', 'shape': (2, 3, 4),
            'spacing': (0.5e-3, 0.4e-3, 0.25e-3), 'num': 2500, 'step': 8e-8, 'flag': True,
            'inner': [['50', '406', 'None'], ['50', '435', 'None']], 'i32': np.arange(5, dtype=np.int32),
            'f64': np.linspace(0, 1, 7)}


def conv(v):
    if isinstance(v, (bytes, np.bytes_)):
        return {'bytes': bytes(v).decode('utf-8', 'replace')}
    if isinstance(v, np.ndarray):
        if v.dtype.kind in 'SO':
            return {'strings': [x.decode() if isinstance(x, bytes) else str(x) for x in v.ravel().tolist()],
                    'shape': list(v.shape)}
        return {'array': v.ravel().tolist(), 'shape': list(v.shape), 'dtype': v.dtype.str}
    if isinstance(v, (np.generic,)):
        return {'scalar': v.item(), 'dtype': v.dtype.str}
    return {'py': v if isinstance(v, (int, float, str, bool)) else str(v)}


def dump(path):
    out = {}
    with h5py.File(path, 'r') as f:
        out['/'] = {'attrs': {k: conv(v) for k, v in f.attrs.items()}}

        def visit(name, obj):
            rec = {'attrs': {k: conv(v) for k, v in obj.attrs.items()}}
            if isinstance(obj, h5py.Dataset):
                rec['data'] = conv(obj[()])
            out[name] = rec
        f.visititems(visit)
    print(json.dumps(out))


def make(path, libver):
    with h5py.File(path, 'w', libver=libver) as f:
        f.attrs['protocol'] = '0.1'
        f.attrs['is_array'] = False
        for k, v in sample().items():
            d = f.create_dataset(k, data=v)
            d.attrs['is_ndarray'] = isinstance(v, np.ndarray)
            d.attrs['is_list'] = isinstance(v, list)
            d.attrs['is_tuple'] = isinstance(v, tuple)
            d.attrs['is_str'] = isinstance(v, str) or (isinstance(v, list) and isinstance(v[0][0], str))
            d.attrs['is_bytes'] = False
            d.attrs['is_none'] = False
        g = f.create_group('space')
        g.attrs['is_array'] = False
        d = g.create_dataset('extra', data=(50, 50))
        for a in ('is_ndarray', 'is_list', 'is_str', 'is_bytes', 'is_none'):
            d.attrs[a] = False
        d.attrs['is_tuple'] = True


if __name__ == '__main__':
    if sys.argv[1] == 'dump':
        dump(sys.argv[2])
    else:
        make(sys.argv[2], sys.argv[3])
