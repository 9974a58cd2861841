"""The C ABI library: it loads and exports every entry point include/sord_amd.h
declares (no device calls)."""
import ctypes
import os
import re

from coseis_amd import solver

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    h = open(os.path.join(ROOT, 'include', 'sord_amd.h')).read()
    return sorted(set(re.findall(r'\b(sord_[a-z_]+)\s*\(', h)))


def test_header_lists_the_wrapper_exports():
    assert sorted(solver.EXPORTS) == declared()


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(solver.LIBPATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_error_paths_without_device():
    lib = solver.load()
    h = ctypes.c_void_p()
    assert lib.sord_create(b'1 2 3', 3, 0, ctypes.byref(h)) == 1  # bad realsize
    assert lib.sord_init(None) == 11
    assert lib.sord_last_error() == b'null handle'
