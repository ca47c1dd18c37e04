"""The C++ host layer's helpers against the reference's own (CPU only): number text
(to_string of float/double, fixed decimals), print_time, converged(), the exported file
names and the Units closed forms (Stokes flow, rectangular duct).  The host layer's side is
fluidx3d_amd/host/bin/libhost_check.so (built by make -C fluidx3d_amd/host); the
reference's side is oracle/_ref/libref_helpers.so, compiled from the reference's sources."""
import ctypes
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def host():
    path = os.path.join(ROOT, "fluidx3d_amd", "host", "bin", "libhost_check.so")
    if not os.path.exists(path):
        pytest.skip("host layer not built (make -C fluidx3d_amd/host)")
    L = ctypes.CDLL(path)
    f32, f64, c = ctypes.c_float, ctypes.c_double, ctypes.c_char_p
    L.host_to_string_float.argtypes = [f32, c, ctypes.c_int]
    L.host_to_string_double.argtypes = [f64, c, ctypes.c_int]
    L.host_to_string_float_decimals.argtypes = [f32, ctypes.c_uint, c, ctypes.c_int]
    L.host_to_string_double_decimals.argtypes = [f64, ctypes.c_uint, c, ctypes.c_int]
    L.host_print_time.argtypes = [f64, c, ctypes.c_int]
    L.host_converged_double.argtypes = [f64] * 4
    L.host_converged_float.argtypes = [f32] * 4
    L.host_default_filename.argtypes = [c, c, c, ctypes.c_ulonglong, c, ctypes.c_int]
    L.host_u_Stokes.argtypes = [f32] * 7 + [ctypes.POINTER(f32)]
    for name, n in (("host_rho_Stokes", 9), ("host_F_Stokes", 4), ("host_f_Stokes", 5), ("host_f_from_u_rectangular_duct", 5),
                    ("host_u_from_Re", 3), ("host_f_from_u_Poiseuille_2D", 4)):
        getattr(L, name).restype = f32
        getattr(L, name).argtypes = [f32] * n
    return L


@pytest.fixture(scope="module")
def ref(refhelpers):
    L = refhelpers
    f32, f64, c = ctypes.c_float, ctypes.c_double, ctypes.c_char_p
    L.ref_to_string_double.argtypes = [f64, c, ctypes.c_int]
    L.ref_to_string_float_decimals.argtypes = [f32, ctypes.c_uint, c, ctypes.c_int]
    L.ref_to_string_double_decimals.argtypes = [f64, ctypes.c_uint, c, ctypes.c_int]
    L.ref_print_time.argtypes = [f64, c, ctypes.c_int]
    L.ref_converged_double.argtypes = [f64] * 4
    L.ref_converged_float.argtypes = [f32] * 4
    L.ref_u_Stokes.argtypes = [f32] * 7 + [ctypes.POINTER(f32)]
    for name, n in (("ref_rho_Stokes", 9), ("ref_f_Stokes", 5), ("ref_f_from_u_rectangular_duct", 5), ("ref_u_from_Re", 3)):
        getattr(L, name).restype = f32
        getattr(L, name).argtypes = [f32] * n
    return L


def text(fn, *args):
    buf = ctypes.create_string_buffer(128)
    n = fn(*args, buf, 128)
    assert n >= 0
    return buf.value.decode()


def sample_floats(count, seed):
    rng = np.random.default_rng(seed)
    bits = rng.integers(0, 2 ** 32, count, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    special = np.array([0.0, -0.0, 1.0, -1.0, 0.1, 9.999999, 10.0, 99.99999, 1e-45, 3.4e38, 1e32, 1e-32, 0.5, 2.5e-7,
                        1.0 / 3.0, 6.0 * np.pi, np.inf, -np.inf, np.nan], dtype=np.float32)
    return np.concatenate([x, special, rng.uniform(-1e4, 1e4, count // 4).astype(np.float32)])


def test_float_and_double_text_like_reference(host, ref):
    for x in sample_floats(20000, 1):
        assert text(host.host_to_string_float, x) == text(ref.ref_to_string_float, x), x
    rng = np.random.default_rng(2)
    doubles = np.concatenate([rng.integers(0, 2 ** 63, 20000, dtype=np.uint64).view(np.float64), rng.uniform(-1e6, 1e6, 5000),
                              np.array([0.0, 1.0, 0.1, 1e256, 1e-300, 123456.789, np.inf, np.nan])])
    for x in doubles:
        assert text(host.host_to_string_double, x) == text(ref.ref_to_string_double, x), x
    # the viscosity line of write_status for nu = 1/6 (the C++ scene test checks this text)
    sixth = np.float32(1.0) / np.float32(6.0)
    assert text(ref.ref_to_string_float, sixth) == text(host.host_to_string_float, sixth) == "1.66666676E-1"


def test_fixed_decimals_like_reference(host, ref):
    rng = np.random.default_rng(3)
    for x in np.concatenate([rng.uniform(-1e5, 1e5, 5000), rng.uniform(-2, 2, 5000), [0.0, 0.5, 0.05, 99.95, 1e19]]):
        for d in (0, 1, 3, 6, 8, 12):
            assert text(host.host_to_string_float_decimals, np.float32(x), d) == text(ref.ref_to_string_float_decimals, np.float32(x), d), (x, d)
            assert text(host.host_to_string_double_decimals, x, d) == text(ref.ref_to_string_double_decimals, x, d), (x, d)


def test_print_time_and_converged_like_reference(host, ref):
    rng = np.random.default_rng(4)
    for t in np.concatenate([[0.0, 0.4, 0.5, 59.4, 59.5, 59.6, 3599.5, 3600, 86399.5, 86400, 31556924.5, 4e8],
                             rng.uniform(0, 1e9, 3000), rng.uniform(0, 5000, 3000)]):
        assert text(host.host_print_time, t) == text(ref.ref_print_time, t), t
    seq = rng.uniform(0.9, 1.1, (5000, 3)) * np.array([1.0, 1.0, 1.0])
    seq[::2] = 1.0 + rng.uniform(-1e-5, 1e-5, (2500, 3))
    for a, b, c in seq:
        for thr in (1e-4, 1e-5):
            assert host.host_converged_double(a, b, c, thr) == ref.ref_converged_double(a, b, c, thr)
            assert host.host_converged_float(a, b, c, thr) == ref.ref_converged_float(a, b, c, thr)


def test_default_filename_like_reference(host):
    """default_filename (src/lbm.cpp:85-89): path prefix (default <exe dir>/export/),
    name-<t as 9 digits>, the extension replacing any after the last '.'."""
    assert text(host.host_default_filename, b"out/", b"u", b".vtk", 1000) == "out/u-000001000.vtk"
    assert text(host.host_default_filename, b"", b"status", b".txt", 12) .endswith("/export/status-000000012.txt")
    assert text(host.host_default_filename, b"a/", b"", b"vtk", 3) == "a/file-000000003.vtk"
    assert text(host.host_default_filename, b"a/", b"rho", b".vtk", 12345678901) == "a/rho-345678901.vtk"


def test_units_closed_forms_like_reference(host, ref, refhelpers):
    rng = np.random.default_rng(5)
    out_h, out_r = (ctypes.c_float * 3)(), (ctypes.c_float * 3)()
    for _ in range(5000):
        p = rng.uniform(-60, 60, 3).astype(np.float32)
        u0 = rng.uniform(-0.01, 0.01, 3).astype(np.float32)
        R = np.float32(rng.uniform(4, 40))
        host.host_u_Stokes(*p, *u0, R, out_h)
        ref.ref_u_Stokes(*p, *u0, R, out_r)
        assert list(out_h) == list(out_r), (p, u0, R)
        assert host.host_rho_Stokes(*p, *u0, R, 1.0, 0.5) == ref.ref_rho_Stokes(*p, *u0, R, 1.0, 0.5)
    for _ in range(500):
        a = rng.uniform(0.1, 3, 5).astype(np.float32)
        assert host.host_f_from_u_rectangular_duct(*a) == ref.ref_f_from_u_rectangular_duct(*a)
        assert host.host_f_Stokes(*a) == ref.ref_f_Stokes(*a)
        assert host.host_F_Stokes(*a[:4]) == refhelpers.ref_F_Stokes(*a[:4])
        assert host.host_u_from_Re(*a[:3]) == ref.ref_u_from_Re(*a[:3])
        assert host.host_f_from_u_Poiseuille_2D(*a[:4]) == refhelpers.ref_f_from_u_Poiseuille_2D(*a[:4])
    # the Stokes scene's velocity (Re = 0.01 on the diameter 64, nu = 1)
    assert host.host_u_from_Re(0.01, 64.0, 1.0) == np.float32(np.float32(0.01) * np.float32(1.0) / np.float32(64.0))
