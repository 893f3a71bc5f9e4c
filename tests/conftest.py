import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs the HIP engine)")


def rel_err(a, b):
    """normwise: max|a - b| / max|b|"""
    a = np.asarray(a)
    b = np.asarray(b)
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / (s if s > 0 else 1.0))


@pytest.fixture(scope="session")
def ref_lib():
    import ctypes
    path = os.path.join(ROOT, "oracle", "_ref", "libdiscamb_ref.so")
    if not os.path.exists(path):
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    lib = ctypes.CDLL(path)
    lib.ref_g_function.restype = ctypes.c_double
    lib.ref_g_function.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
    return lib


@pytest.fixture(scope="session")
def oracle_lib():
    import ctypes
    import hc_oracle
    lib = hc_oracle.lib()
    lib.hc_oracle_g.restype = ctypes.c_double
    lib.hc_oracle_g.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
    return lib
