"""The oracle's attachment clear (sw::Blitter::clear / fastClear,
src/Device/Blitter.cpp:55-326) against closed forms.

RGBA8 packing is pinned by the exact codes k/255 -> k, by its distance to
the real 255 c (at most one half, so the packing is a rounding to nearest),
by monotonicity, and by the clamp's special cases; D32F, S8 and RGBA32F by
bit patterns that a conversion or a clamp would change.  D16 (the blit
routine from a D32F pixel) is pinned the same way as RGBA8, plus the NaN
case the ordered Min decides; B8G8R8A8 is R8G8B8A8 with R and B exchanged.
"""
import numpy as np

from oracle_runner import oracle_clear
from swiftshader_amd import abi


def _rgba8(c):
    t = np.full((1, 1, 1, 4), 0x5A, np.uint8)
    oracle_clear(t, abi.FORMAT_R8G8B8A8_UNORM, [c, c, c, c])
    assert (t[0, 0, 0] == t[0, 0, 0, 0]).all()
    return int(t[0, 0, 0, 0])


def test_unorm8_exact_codes():
    for k in range(256):
        assert _rgba8(np.float32(k) / np.float32(255)) == k


def test_unorm8_rounds_to_nearest_and_is_monotone():
    xs = np.linspace(0.0, 1.0, 20001, dtype=np.float32)
    got = np.array([_rgba8(x) for x in xs])
    assert np.all(np.diff(got) >= 0)
    assert np.max(np.abs(got - 255.0 * xs.astype(np.float64))) <= 0.5 + 1e-4


def test_unorm8_clamp_and_special_values():
    assert _rgba8(-0.3) == 0
    assert _rgba8(1.7) == 255
    assert _rgba8(float("-inf")) == 0
    assert _rgba8(float("inf")) == 255
    assert _rgba8(float("nan")) == 0
    assert _rgba8(-0.0) == 0


def test_unorm8_channel_order():
    t = np.zeros((1, 1, 1, 4), np.uint8)
    oracle_clear(t, abi.FORMAT_R8G8B8A8_UNORM, [1.0, 0.0, 128 / 255, 64 / 255])
    assert list(t[0, 0, 0]) == [255, 0, 128, 64]


def test_depth_is_written_as_given():
    # D32F is not a normalised format: no clamp, the float's bits as they are.
    for bits in (0x40200000, 0xBF800000, 0x7FC00001, 0x00000001, 0x80000000):
        v = np.array([bits], np.uint32).view(np.float32)[0]
        t = np.zeros((1, 2, 3), np.float32)
        oracle_clear(t, abi.FORMAT_D32_SFLOAT, v)
        assert (t.view(np.uint32) == bits).all(), hex(bits)


def test_stencil_takes_the_low_byte():
    for v, want in ((0x1FF, 0xFF), (0x80, 0x80), (0x100, 0), (7, 7)):
        t = np.full((1, 2, 2), 0x33, np.uint8)
        oracle_clear(t, abi.FORMAT_S8_UINT, v)
        assert (t == want).all()


def test_float_colour_is_copied_bit_for_bit():
    bits = np.array([0x7FC00123, 0x80000000, 0x00000005, 0x4B7FFFFF], np.uint32)
    t = np.zeros((4, 3, 4, 4), np.float32)
    oracle_clear(t, abi.FORMAT_R32G32B32A32_SFLOAT, bits.view(np.float32))
    assert (t.view(np.uint32) == bits).all()


def test_area_is_clipped_and_covers_every_sample():
    rng = np.random.default_rng(3)
    t = rng.integers(0, 255, (4, 9, 11), dtype=np.uint8)
    before = t.copy()
    oracle_clear(t, abi.FORMAT_S8_UINT, 200, area=(-3, 5, 8, 10))
    inside = np.zeros((9, 11), bool)
    inside[5:9, 0:5] = True
    for q in range(4):
        assert (t[q][inside] == 200).all()
        assert (t[q][~inside] == before[q][~inside]).all()


def _d16(v):
    t = np.full((1, 1, 2), 0x1234, np.uint16)
    oracle_clear(t, abi.FORMAT_D16_UNORM, v)
    assert t[0, 0, 0] == t[0, 0, 1]
    return int(t[0, 0, 0])


def test_d16_exact_codes_and_rounding():
    # k / 65535 -> k for every code; the packing is within one half of
    # 65535 d (plus the float product's half ulp, 2^-9 near 65535) and
    # monotone: a rounding to nearest.
    ks = np.arange(65536)
    got = np.array([_d16(np.float32(k) / np.float32(65535)) for k in ks[::97]] + [_d16(1.0)])
    assert np.array_equal(got, np.append(ks[::97], 65535))
    xs = np.linspace(0.0, 1.0, 5001, dtype=np.float32)
    got = np.array([_d16(x) for x in xs])
    assert np.all(np.diff(got) >= 0)
    assert np.max(np.abs(got - 65535.0 * xs.astype(np.float64))) <= 0.5 + 2.0 ** -8


def test_d16_clamp_and_special_values():
    assert _d16(-0.25) == 0
    assert _d16(3.0) == 65535
    assert _d16(float("-inf")) == 0
    assert _d16(float("inf")) == 65535
    # A NaN passes sw::clamp, then Min(NaN, 0xFFFF) selects 0xFFFF.
    assert _d16(float("nan")) == 65535


def test_bgra8_is_rgba8_with_red_and_blue_exchanged():
    rng = np.random.default_rng(4)
    for _ in range(50):
        c = rng.uniform(-0.2, 1.2, 4).astype(np.float32)
        a = np.zeros((1, 1, 1, 4), np.uint8)
        b = np.zeros((1, 1, 1, 4), np.uint8)
        oracle_clear(a, abi.FORMAT_R8G8B8A8_UNORM, c)
        oracle_clear(b, abi.FORMAT_B8G8R8A8_UNORM, c)
        assert list(b[0, 0, 0]) == [a[0, 0, 0, 2], a[0, 0, 0, 1], a[0, 0, 0, 0], a[0, 0, 0, 3]]
