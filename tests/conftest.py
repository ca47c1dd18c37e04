import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library through the C ABI)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def refhelpers():
    """oracle/_ref/libref_helpers.so built from the reference's own host sources (skips where absent)."""
    import ctypes
    path = os.path.join(ROOT, "oracle", "_ref", "libref_helpers.so")
    if not os.path.exists(path):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    L = ctypes.CDLL(path)
    L.ref_to_string_float.argtypes = [ctypes.c_float, ctypes.c_char_p, ctypes.c_int]
    L.ref_tau.restype = ctypes.c_float
    L.ref_tau.argtypes = [ctypes.c_float]
    L.ref_half_to_float_custom.restype = ctypes.c_float
    L.ref_half_to_float_custom.argtypes = [ctypes.c_ushort]
    L.ref_float_to_half_custom.restype = ctypes.c_ushort
    L.ref_float_to_half_custom.argtypes = [ctypes.c_float]
    L.ref_half_to_float.restype = ctypes.c_float
    L.ref_half_to_float.argtypes = [ctypes.c_ushort]
    L.ref_random.restype = ctypes.c_float
    L.ref_random.argtypes = [ctypes.POINTER(ctypes.c_uint), ctypes.c_float]
    L.ref_random_symmetric.restype = ctypes.c_float
    L.ref_random_symmetric.argtypes = [ctypes.POINTER(ctypes.c_uint), ctypes.c_float]
    L.ref_sphere.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float]
    L.ref_f_from_u_Poiseuille_2D.restype = ctypes.c_float
    L.ref_f_from_u_Poiseuille_2D.argtypes = [ctypes.c_float] * 4
    L.ref_F_Stokes.restype = ctypes.c_float
    L.ref_F_Stokes.argtypes = [ctypes.c_float] * 4
    return L
