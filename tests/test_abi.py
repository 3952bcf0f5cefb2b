"""The C-ABI library loads without a GPU and exports every symbol the header declares."""

import ctypes
import os
import re

import pytest

from stride_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, 'include', 'stride_amd.h')).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(sa_[a-z_0-9]+)\s*\(', text)))


def test_header_declares_binding_exports():
    assert set(_declared()) == set(_native.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = _native.lib()
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.sa_version().decode().startswith('stride_amd')


def test_errors_are_reported_not_crashed():
    lib = _native.lib()
    out = ctypes.c_void_p()
    # null grid -> argument error with a message
    assert lib.sa_create(None, ctypes.byref(out)) == -1
    assert b'null' in lib.sa_last_error()
    g = _native._Grid()
    g.dim = 4
    assert lib.sa_create(ctypes.byref(g), ctypes.byref(out)) == -1
    assert b'dim' in lib.sa_last_error()
    # calls on a null solver fail cleanly
    assert lib.sa_forward(None, None, None) == -1
    assert lib.sa_adjoint(None, None, None, None, None, None, None) == -1


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_native, '_lib', None)
    monkeypatch.setattr(_native, 'library_path', '/nonexistent/libstride_amd.so')
    with pytest.raises(_native.NativeLibraryError):
        _native.lib()
