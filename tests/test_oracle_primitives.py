"""Pins the oracle's numeric primitives and constant tables.

Known answers come from the reference's own unit tests
(tests/ReactorUnitTests/ReactorUnitTests.cpp) and from the compiled
reference sources in oracle/_ref (src/Pipeline/Constants.cpp); the rest
is pinned by the mathematics each primitive implements.
"""
import ctypes
import math
import struct

import numpy as np
import pytest

from oracle_runner import oracle, oracle_render, reference_lib


def f32(x):
    return struct.unpack("f", struct.pack("f", x))[0]


def test_round_int_known_answers():
    # ReactorUnitTests.cpp:1495-1530 (RoundInt / RoundIntClamped on x86).
    L = oracle()
    assert [L.swo_round_int(v) for v in (3.1, 3.6, -3.1, -3.6)] == [3, 4, -3, -4]
    assert L.swo_round_int_clamped(2147483648.0) >= 2147483520
    assert L.swo_round_int_clamped(-2147483648.0) < -2147483647
    assert L.swo_round_int_clamped(2147483520.0) == 2147483520
    assert L.swo_round_int_clamped(-2147483520.0) == -2147483520


def test_round_int_is_half_even_and_saturates_to_int_min():
    L = oracle()
    # IEEE round-to-nearest-even (cvtps2dq).
    assert [L.swo_round_int(v) for v in (0.5, 1.5, 2.5, -0.5, -1.5)] == [0, 2, 2, 0, -2]
    # Out of range and NaN give 0x80000000.
    for v in (float("nan"), 3e9, -3e9, float("inf")):
        assert L.swo_round_int(v) == -2147483648
    # RoundIntClamped maps NaN through Min(x, 2147483520) -> 2147483520.
    assert L.swo_round_int_clamped(float("nan")) == 2147483520


def test_trunc_int():
    L = oracle()
    assert [L.swo_trunc_int(v) for v in (3.9, -3.9, 0.999, -0.999)] == [3, -3, 0, 0]
    for v in (float("nan"), 2147483648.0, -2147483904.0):
        assert L.swo_trunc_int(v) == -2147483648


def test_mulhigh_known_answers():
    # ReactorUnitTests.cpp:1813-1864: lanes (0x01AA,0x02DD,0x03EE,0xF422) x (0x01BB,0x02CC,0x03FF,0xF411).
    L = oracle()
    a = [0x01AA, 0x02DD, 0x03EE, 0xF422]
    b = [0x01BB, 0x02CC, 0x03FF, 0xF411]
    pack = lambda x, y: x | (y << 16)
    s0 = L.swo_mulhigh_s16x2(pack(a[0], a[1]), pack(b[0], b[1]))
    s1 = L.swo_mulhigh_s16x2(pack(a[2], a[3]), pack(b[2], b[3]))
    u0 = L.swo_mulhigh_u16x2(pack(a[0], a[1]), pack(b[0], b[1]))
    u1 = L.swo_mulhigh_u16x2(pack(a[2], a[3]), pack(b[2], b[3]))
    assert (s0, s1) == (0x00080002, 0x008D000F)
    assert (u0, u1) == (0x00080002, 0xE8C0000F)


def test_log2sqrt_bound_and_exact_points():
    # log2sqrt(L) = 0.25 * (bits(L*L) - bits(1.0)) * 2^-23: a piecewise-linear
    # under-estimate of log2(sqrt(L)) by at most max(log2(1+m) - m)/4 = 0.0215,
    # exact where L*L is a power of two.
    L = oracle()
    rng = np.random.default_rng(0)
    xs = np.exp(rng.uniform(np.log(1e-6), np.log(1e8), 4000)).astype(np.float32)
    for x in xs:
        approx = L.swo_log2sqrt(float(x))
        true = 0.5 * math.log2(float(x))
        assert -2e-5 <= true - approx <= 0.02152 + 2e-5, (x, approx, true)
    for k in range(-20, 21):
        x = 2.0 ** (k / 2.0) if k % 2 == 0 else 2.0 ** (k // 2)
        x = f32(x)
        if (x * x) == 2.0 ** round(math.log2(x * x)):
            assert L.swo_log2sqrt(x) == pytest.approx(0.5 * math.log2(x), abs=1e-6)


def test_frac_clamped_below_one():
    L = oracle()
    assert L.swo_frac(2.25) == 0.25
    assert L.swo_frac(-0.25) == 0.75
    # x - floor(x) rounds to 1.0 for tiny negative x; Frac clamps to 0x3F7FFFFF.
    assert L.swo_frac(-1e-9) == struct.unpack("f", struct.pack("I", 0x3F7FFFFF))[0]


def _oracle_constants():
    L = oracle()
    xf = (ctypes.c_int * 4)()
    yf = (ctypes.c_int * 4)()
    ms = (ctypes.c_int * 2)()
    cw = (ctypes.c_uint16 * 17)()
    uw = (ctypes.c_float * 17)()
    us = (ctypes.c_float * 17)()
    L.swo_constants(xf, yf, ms, cw, uw, us)
    return list(xf), list(yf), list(ms), list(cw), list(uw), list(us)


def test_constants_known_values():
    xf, yf, ms, cw, uw, us = _oracle_constants()
    # Standard 4x locations (Vulkan spec) in 28.4 relative to the centre.
    assert xf == [-2, 6, -6, 2]
    assert yf == [-6, -2, 2, 6]
    assert ms == [9, 21]
    assert cw[:5] == [0xFFFF, 0xFFFF, 0x8000, 0x5555, 0x4000] and cw[16] == 0x1000
    for n in range(1, 17):
        assert uw[n] == f32(1.0 / n)
        assert us[n] == pytest.approx(-(n - 1) / (2 * n), abs=1e-7)


def test_constants_match_compiled_reference():
    ref = reference_lib()
    if ref is None:
        pytest.skip("oracle/_ref/libswr_ref.so not built (needs the reference sources)")
    xf = (ctypes.c_int * 4)()
    yf = (ctypes.c_int * 4)()
    ms = (ctypes.c_int * 2)()
    cw = (ctypes.c_uint16 * 17)()
    uw = (ctypes.c_float * 17)()
    us = (ctypes.c_float * 17)()
    ref.ref_constants(xf, yf, ms, cw, uw, us)
    mine = _oracle_constants()
    assert list(xf) == mine[0] and list(yf) == mine[1] and list(ms) == mine[2]
    assert list(cw) == mine[3]
    assert [struct.pack("f", v) for v in uw] == [struct.pack("f", v) for v in mine[4]]
    assert [struct.pack("f", v) for v in us] == [struct.pack("f", v) for v in mine[5]]


def test_constants_match_golden_vectors():
    """The committed tables of the compiled reference (tests/golden/make_ref_vectors.py)."""
    import os
    import numpy as np
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "constants_ref.npz"))
    xf, yf, ms, cw, uw, us = _oracle_constants()
    assert xf == list(z["xf"]) and yf == list(z["yf"]) and ms == list(z["ms_offsets"])
    assert cw == list(z["c_weight"])
    assert np.array_equal(np.array(uw, np.float32).view(np.uint32), z["uv_weight"].view(np.uint32))
    assert np.array_equal(np.array(us, np.float32).view(np.uint32), z["uv_start"].view(np.uint32))


def test_index_types_read_the_same_indices():
    """UINT32, UINT16 and UINT8_EXT index buffers holding the same values
    (Renderer.cpp:696-716) render the same image bit for bit."""
    from cases import _indexed, _strip_band
    for make in (lambda: _indexed(seed=57), lambda: _strip_band(seed=58)):
        out = []
        for dt in (np.uint32, np.uint16, np.uint8):
            d = make()
            d.indices = d.indices.astype(dt)
            out.append(oracle_render(d))
        for o in out[1:]:
            assert np.array_equal(o[0], out[0][0]) and np.array_equal(o[1].view(np.uint32), out[0][1].view(np.uint32))
            assert o[3] == out[0][3]
