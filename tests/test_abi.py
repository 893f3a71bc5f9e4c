"""CPU checks of the boundary: the C ABI library loads and exports every function that
include/discamb_amd.h declares; error paths that need no GPU."""
import ctypes
import os
import re

import pytest

from discamb_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "discamb_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", " ", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dh_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 14
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES)


def test_version_and_error_without_gpu():
    lib = _lib.load()
    assert b"gfx950" in lib.dh_version()
    n = ctypes.c_int(-1)
    assert lib.dh_device_count(ctypes.byref(n)) == 0
    assert n.value >= 0
    # create with a null description fails cleanly with an argument error
    h = ctypes.c_void_p()
    rc = lib.dh_hc_create(None, 0, ctypes.byref(h))
    assert rc == 1
    assert b"null" in lib.dh_last_error()


def test_calls_on_null_calculator_fail_cleanly():
    lib = _lib.load()
    assert lib.dh_hc_set_calculation_mode(None, b"HIP") == 1
    assert lib.dh_hc_run(None, 7) == 1
