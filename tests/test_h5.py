"""Mosaic's HDF5 protocol 0.1 without h5py (stride_amd.h5): the lookup3 checksum's published
values, round trips of every value kind, the reference's breast2D fixture, and (where an
interpreter with h5py exists on this machine) files read and written by h5py in both
directions and both format generations."""

import json
import os
import subprocess

import numpy as np
import pytest

from stride_amd import h5

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = '/root/reference/stride_examples/examples/breast2D/data/anastasio2D-TrueModel.h5'
H5PY_PYTHON = '/opt/conda/bin/python3.9'


def _h5py_available():
    if not os.path.exists(H5PY_PYTHON):
        return False
    return subprocess.run([H5PY_PYTHON, '-c', 'import h5py'], capture_output=True).returncode == 0


def test_lookup3_published_values():
    # lookup3.c (Jenkins, 2006) driver values for hashlittle
    assert h5.lookup3(b'') == 0xDEADBEEF
    assert h5.lookup3(b'Four score and seven years ago') == 0x17770551
    assert h5.lookup3(b'Four score and seven years ago', 1) == 0xCD628161


def _sample():
    return {'data': np.arange(60, dtype=np.float32).reshape(3, 4, 5) / 7, 'name': 'vp', 'shape': (3, 4, 5),
            'spacing': (0.5e-3, 0.4e-3, 0.2e-3), 'inner': [['50', '406', 'None'], ['50', '435', 'None']],
            'none': None, 'flag': True, 'num': 2500, 'step': 8e-8, 'blob': b'\x00\x01\xffabc',
            'i32': np.arange(-3, 4, dtype=np.int32), 'u8': np.arange(5, dtype=np.uint8), 'f64': np.linspace(0, 1, 9),
            'empty': np.zeros((0, 3), np.float32), 'b': np.array([True, False, True]),
            'space': {'extra': (50, 50), 'absorbing': (40, 40), 'time': {'start': 0.0, 'num': 10}},
            'shots': [{'id': k, 'coords': np.full((2, 3), k, np.float64)} for k in range(3)],
            'text': 'utf-8: µs'}


def _check_equal(got, want):
    assert type(got) is type(want) or (isinstance(want, np.ndarray) and isinstance(got, np.ndarray)), (got, want)
    if isinstance(want, dict):
        assert sorted(got) == sorted(want)
        for k in want:
            _check_equal(got[k], want[k])
    elif isinstance(want, list) and want and isinstance(want[0], dict):
        assert len(got) == len(want)
        for g, w in zip(got, want):
            _check_equal(g, w)
    elif isinstance(want, np.ndarray):
        assert got.dtype == want.dtype and got.shape == want.shape
        np.testing.assert_array_equal(got, want)
    else:
        assert got == want


def test_roundtrip(tmp_path):
    f = str(tmp_path / 'a.h5')
    h5.write(f, _sample())
    _check_equal(h5.read(f), _sample())


def test_many_links_and_versions(tmp_path):
    desc = {'k%02d' % i: np.full(4, i, np.float32) for i in range(40)}
    with h5.HDF5(project_name='proj', parameter='vp_grad', path=str(tmp_path), mode='w') as f:
        f.dump(desc)
        assert os.path.basename(f.filename) == 'proj-vpGrad.h5'
    with h5.HDF5(project_name='proj', parameter='vp_grad', path=str(tmp_path), mode='w') as f:
        f.dump({'x': 1})
        assert os.path.basename(f.filename) == 'proj-vpGrad-00001.h5'
    with h5.HDF5(project_name='proj', parameter='vp_grad', path=str(tmp_path), mode='r') as f:
        assert f.load() == {'x': 1}  # the last version
    assert h5.file_exists(project_name='proj', parameter='vp_grad', version=1, path=str(tmp_path))
    _check_equal(h5.read(str(tmp_path / 'proj-vpGrad.h5')), desc)


@pytest.mark.skipif(not os.path.exists(FIXTURE), reason='reference tree not present')
def test_breast2d_fixture():
    d = h5.read(FIXTURE)
    fx = np.load(os.path.join(HERE, 'golden', 'anastasio2D_vp.npz'))
    np.testing.assert_array_equal(d['data'], fx['vp'])
    assert d['shape'] == (356, 385) and d['extended_shape'] == (456, 485)
    assert d['space'] == {'absorbing': (40, 40), 'extra': (50, 50), 'shape': (356, 385), 'spacing': (0.0005, 0.0005)}
    assert d['time']['num'] == 2500 and d['time']['step'] == 8e-08
    assert d['inner'] == [['50', '406', 'None'], ['50', '435', 'None']]
    assert d['dtype'] == 'float32' and d['time_dependent'] is False


@pytest.mark.skipif(not _h5py_available(), reason='no interpreter with h5py on this machine')
def test_h5py_reads_our_files(tmp_path):
    f = str(tmp_path / 'ours.h5')
    h5.write(f, _sample())
    out = subprocess.run([H5PY_PYTHON, os.path.join(HERE, 'h5py_check.py'), 'dump', f], capture_output=True,
                         text=True, check=True)
    seen = json.loads(out.stdout)
    s = _sample()
    assert seen['data']['data']['shape'] == [3, 4, 5]
    np.testing.assert_array_equal(np.array(seen['data']['data']['array'], np.float32), s['data'].ravel())
    assert seen['num']['data'] == {'scalar': 2500, 'dtype': '<i8'}
    assert seen['flag']['data'] == {'scalar': True, 'dtype': '|b1'}
    assert seen['name']['data'] == {'bytes': 'vp'}
    assert seen['inner']['data']['strings'] == ['50', '406', 'None', '50', '435', 'None']
    assert seen['space/time/start']['data']['scalar'] == 0.0
    assert seen['shots/shots_00000002/coords']['data']['array'] == [2.0] * 6
    assert seen['/']['attrs']['protocol'] == {'bytes': '0.1'}
    assert seen['data']['attrs']['is_ndarray'] == {'scalar': True, 'dtype': '|b1'}


@pytest.mark.skipif(not _h5py_available(), reason='no interpreter with h5py on this machine')
@pytest.mark.parametrize('libver', ['latest', 'earliest'])
def test_we_read_h5py_files(tmp_path, libver):
    f = str(tmp_path / ('h5py_%s.h5' % libver))
    subprocess.run([H5PY_PYTHON, os.path.join(HERE, 'h5py_check.py'), 'make', f, libver], check=True)
    d = h5.read(f)
    np.testing.assert_array_equal(d['data'], np.arange(24, dtype=np.float32).reshape(2, 3, 4) / 7)
    assert d['name'] == 'vp' and d['shape'] == (2, 3, 4) and d['spacing'] == (0.5e-3, 0.4e-3, 0.25e-3)
    assert d['num'] == 2500 and d['step'] == 8e-8 and d['flag'] is True
    assert d['inner'] == [['50', '406', 'None'], ['50', '435', 'None']]
    np.testing.assert_array_equal(d['i32'], np.arange(5, dtype=np.int32))
    assert d['space'] == {'extra': (50, 50)}


def test_scalar_field_and_traces_dump_load(tmp_path):
    from stride_amd.problem import Space, ScalarField, Traces
    space = Space(shape=(30, 40), spacing=0.5e-3, extra=(5, 6), absorbing=(4, 4))
    rng = np.random.default_rng(0)
    vp = ScalarField('vp', space=space, data=(1500 + rng.standard_normal(space.shape)).astype(np.float32))
    fname = vp.dump(path=str(tmp_path), project_name='run')
    assert os.path.basename(fname) == 'run-vp.h5'
    back = ScalarField('vp').load(path=str(tmp_path), project_name='run')
    np.testing.assert_array_equal(back.extended_data, vp.extended_data)  # inner data, edge-padded again
    assert back.space.shape == space.shape and back.space.extra == space.extra
    desc = h5.read(fname)
    assert desc['inner'] == [['5', '35', 'None'], ['6', '46', 'None']] and desc['dtype'] == 'float32'
    tr = Traces('observed', data=rng.standard_normal((7, 50)).astype(np.float32), transducer_ids=np.arange(7) + 3)
    tr.dump(path=str(tmp_path), project_name='run', version=2)
    tb = Traces('observed').load(path=str(tmp_path), project_name='run', version=2)
    np.testing.assert_array_equal(tb.data, tr.data)
    np.testing.assert_array_equal(tb.transducer_ids, np.arange(7) + 3)


@pytest.mark.skipif(not os.path.exists(FIXTURE), reason='reference tree not present')
def test_breast2d_fixture_as_scalar_field(tmp_path):
    from stride_amd.problem import ScalarField
    vp = ScalarField('vp').load(filename=FIXTURE)
    assert vp.extended_shape == (456, 485) and vp.shape == (356, 385)
    np.testing.assert_array_equal(vp.data, np.load(os.path.join(HERE, 'golden', 'anastasio2D_vp.npz'))['vp'])
    assert vp.space.extra == (50, 50) and vp.space.absorbing == (40, 40)
