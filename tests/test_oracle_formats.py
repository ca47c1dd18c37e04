"""Pins the oracle's storage formats and relaxation-rate formatting (CPU only).

FP16S (src/lbm.cpp:410-414): IEEE binary16 RNE, checked against the x86 F16C
conversions for every binary32 input and against numpy.float16.  FP16C
(src/kernel.cpp:848-859): against the closed-form 1-4-11 value of every code, the
reference's own host helpers (oracle/_ref, built from src/utilities.hpp), and the
rounding rule (nearest, ties away from zero).  def_w (src/lbm.cpp:367): against the
reference's own to_string(float).
"""
import ctypes

import numpy as np
import pytest


def test_f16_software_matches_f16c_for_every_float(oracle):
    assert oracle.lib().fxo_f16_selfcheck(8) == 0


def test_f16_matches_numpy(oracle):
    rng = np.random.default_rng(0)
    bits = rng.integers(0, 2 ** 32, 2 ** 22, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    x = x[~np.isnan(x)]
    # plus exact ties between adjacent halves (mantissa bit 12 set, rest zero)
    ties = (np.arange(0x3000, 0x4700, dtype=np.uint32) << 13 | 0x1000).view(np.float32)
    x = np.concatenate([x, ties, -ties, np.float32([0.0, -0.0, 65504.0, 65520.0, 6.0e-8, 2.9802322e-8])])
    with np.errstate(over="ignore"):
        expect = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(oracle.encode("F16", x), expect)
    codes = np.arange(65536, dtype=np.uint16)
    dec = oracle.decode("F16", codes)
    ref = codes.view(np.float16).astype(np.float32)
    same = (dec == ref) | (np.isnan(dec) & np.isnan(ref))
    assert same.all()


def test_fp16s_is_scaled_binary16(oracle):
    rng = np.random.default_rng(1)
    x = (rng.uniform(-1, 1, 200000) * 10.0 ** rng.uniform(-9, 0.2, 200000)).astype(np.float32)
    c = oracle.encode("FP16S", x)
    assert np.array_equal(c, (x * np.float32(32768.0)).astype(np.float16).view(np.uint16))
    codes = np.arange(65536, dtype=np.uint16)
    d = oracle.decode("FP16S", codes)
    with np.errstate(invalid="ignore"):
        ref = codes.view(np.float16).astype(np.float32) * np.float32(2.0 ** -15)
    assert np.array_equal(d[np.isfinite(ref)], ref[np.isfinite(ref)])


def fp16c_value(code):
    """Closed-form value of a 1-4-11 FP16C code, exponent bias 15, computed in float64."""
    code = np.asarray(code, dtype=np.int64)
    s = np.where(code & 0x8000, -1.0, 1.0)
    e = (code >> 11) & 0xF
    m = code & 0x7FF
    v = np.where(e > 0, np.ldexp(1.0 + m / 2048.0, e - 15), np.ldexp(m.astype(np.float64), -25))
    return s * v


def test_fp16c_decode_closed_form(oracle):
    codes = np.arange(65536, dtype=np.uint16)
    d = oracle.decode("FP16C", codes).astype(np.float64)
    assert np.array_equal(d, fp16c_value(codes))


def test_fp16c_roundtrip_every_code(oracle):
    codes = np.arange(65536, dtype=np.uint16)
    back = oracle.encode("FP16C", oracle.decode("FP16C", codes))
    assert np.array_equal(back, codes)


def test_fp16c_rounds_to_nearest_ties_away(oracle):
    # every pair of adjacent positive normal codes: midpoint rounds up (away from zero),
    # just below rounds down, just above rounds up
    k = np.arange(0x0800, 0x7FFE, dtype=np.int64)
    lo, hi = fp16c_value(k), fp16c_value(k + 1)
    mid = ((lo + hi) / 2).astype(np.float32)
    assert np.array_equal(mid.astype(np.float64), (lo + hi) / 2)  # midpoints are exact in binary32
    assert np.array_equal(oracle.encode("FP16C", mid), (k + 1).astype(np.uint16))
    below = np.nextafter(mid, np.float32(0))
    above = np.nextafter(mid, np.float32(4))
    assert np.array_equal(oracle.encode("FP16C", below), k.astype(np.uint16))
    assert np.array_equal(oracle.encode("FP16C", above), (k + 1).astype(np.uint16))
    assert np.array_equal(oracle.encode("FP16C", -mid), (k + 1 | 0x8000).astype(np.uint16))
    # random values: the decoded result is within half a code step
    rng = np.random.default_rng(2)
    x = (rng.uniform(0.5, 1.0, 100000) * 2.0 ** rng.integers(-13, 1, 100000)).astype(np.float32)
    d = oracle.decode("FP16C", oracle.encode("FP16C", x)).astype(np.float64)
    ulp = np.ldexp(1.0, np.floor(np.log2(x.astype(np.float64))).astype(int) - 11)
    assert np.all(np.abs(d - x) <= ulp / 2)


def test_fp16c_matches_reference_host_helpers(oracle, refhelpers):
    """src/utilities.hpp:133-144 (host copies of the device codec).  The host encoder only
    adds saturation for |x| >= 2 after rounding, so inputs stay below 1.999."""
    codes = np.arange(65536, dtype=np.uint16)
    ours = oracle.decode("FP16C", codes)
    ref = np.array([refhelpers.ref_half_to_float_custom(int(c)) for c in codes], dtype=np.float32)
    assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))
    rng = np.random.default_rng(4)
    x = (rng.uniform(-1.999, 1.999, 30000) * 10.0 ** rng.uniform(-10, 0, 30000)).astype(np.float32)
    ref_codes = np.array([refhelpers.ref_float_to_half_custom(float(v)) for v in x], dtype=np.uint16)
    assert np.array_equal(oracle.encode("FP16C", x), ref_codes)


def ref_text(refhelpers, x):
    buf = ctypes.create_string_buffer(64)
    n = refhelpers.ref_to_string_float(ctypes.c_float(x), buf, 64)
    assert n > 0
    return buf.value.decode()


def test_float_text_matches_reference(oracle, refhelpers):
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.uniform(0.0, 2.0, 3000), 10.0 ** rng.uniform(-20, 20, 2000), [1.0, 0.1, 1e-7, 9.99999999]]).astype(np.float32)
    for x in xs:
        assert oracle.format_float(float(x)) == ref_text(refhelpers, float(x)), x


@pytest.mark.parametrize("nu", [1.0 / 6.0, 0.048, 0.01, 0.02, 0.0125, 1.0, 0.5, 0.001, 3.0e-4, 0.1234567])
def test_relaxation_rate_matches_reference(oracle, refhelpers, nu):
    """w = float literal the reference prints for 1/tau (src/lbm.cpp:367), re-parsed."""
    import fluidx3d_amd as fx
    tau = refhelpers.ref_tau(nu)
    text = ref_text(refhelpers, np.float32(1.0) / np.float32(tau))
    expect = np.float32(float(text))  # correctly rounded parse, like the device compiler
    assert oracle.relaxation_rate(nu) == expect
    assert fx.relaxation_rate(nu) == expect  # the product's host-side restatement


def test_relaxation_rate_random_nu(oracle, refhelpers):
    import fluidx3d_amd as fx
    rng = np.random.default_rng(6)
    for nu in (10.0 ** rng.uniform(-4, 0.5, 400)).astype(np.float32):
        tau = refhelpers.ref_tau(float(nu))
        expect = np.float32(float(ref_text(refhelpers, np.float32(1.0) / np.float32(tau))))
        assert fx.relaxation_rate(float(nu)) == expect == oracle.relaxation_rate(float(nu))


def test_lcg_and_sphere_helpers(refhelpers):
    """Scene helpers used by the tests follow the reference's LCG and sphere()."""
    from tests import scenes
    seed = ctypes.c_uint(42)
    s = 42
    for _ in range(1000):
        r = refhelpers.ref_random(ctypes.byref(seed), ctypes.c_float(1.0))
        s = (1103515245 * s + 12345) % 2147483648
        assert np.float32(r) == np.float32(np.float32(1.0) * np.float32(s) * np.float32(4.65661287e-10))
    flags = scenes.sphere_flags(20, 18, 16, (9.5, 8.0, 7.25), 5.5)
    n = 0
    for z in range(16):
        for y in range(18):
            for x in range(20):
                assert bool(flags[n]) == bool(refhelpers.ref_sphere(x, y, z, 9.5, 8.0, 7.25, 5.5))
                n += 1
