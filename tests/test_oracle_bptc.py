"""The oracle's BC6H and BC7 decoders (oracle/swr_oracle_bptc.c).

* bit for bit against the reference's BC_Decoder.cpp compiled unmodified
  into oracle/_ref (when the reference sources are present): random blocks
  forced into each BC7 mode (0-7 and the invalid all-zero mode byte) and
  each BC6H mode (the 14 defined ones and the reserved codes), signed and
  unsigned, on ragged image sizes, bytes outside the image untouched;
* closed forms from the format definitions: a BC7 mode-6 block with equal
  endpoints is that colour everywhere, whatever its indices; BC7 mode 6
  with endpoints 0 and 255 gives the 4-bit weights w / 64 * 255 rounded;
  BC6H mode 3 (unsigned) from endpoint 0 to the largest 10-bit value spans
  0 to 0x7BFF (65504, the largest half float) with the 4-bit weights; an
  invalid BC7 block is transparent black and an invalid BC6H block
  (0, 0, 0, 1.0).
"""
import numpy as np
import pytest

from oracle_runner import oracle_decode_bc, reference_lib

BC6_MODES = [0, 1, 2, 6, 10, 14, 18, 22, 26, 30, 3, 7, 11, 15]
BC6_RESERVED = [19, 23, 27, 31]


def _decode(blocks, w, h, n, flag, lib=None):
    bpp = 8 if n == 6 else 4
    pitch = w * bpp + 8
    dst = np.full((h, pitch), 0xA5, np.uint8)
    assert oracle_decode_bc(blocks, dst, w, h, pitch, bpp, n, flag, lib) == 1
    return dst, bpp


def _bc7_blocks(rng, nb, mode):
    b = rng.integers(0, 256, (nb, 16), dtype=np.uint8)
    if mode == 8:
        b[:, 0] = 0
    else:
        keep = np.uint8((0xFF << (mode + 1)) & 0xFF)
        b[:, 0] = (b[:, 0] & keep) | np.uint8(1 << mode)
    return b.ravel()


def _bc6_blocks(rng, nb, mode):
    b = rng.integers(0, 256, (nb, 16), dtype=np.uint8)
    if mode in (0, 1):
        b[:, 0] = (b[:, 0] & np.uint8(0xFC)) | np.uint8(mode)
    else:
        b[:, 0] = (b[:, 0] & np.uint8(0xE0)) | np.uint8(mode)
    return b.ravel()


@pytest.mark.parametrize("mode", range(9))
def test_bc7_matches_reference(mode):
    ref = reference_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (no reference sources)")
    rng = np.random.default_rng(70 + mode)
    for w, h in ((4, 4), (9, 6), (32, 16), (1, 3)):
        nb = ((w + 3) // 4) * ((h + 3) // 4)
        blocks = _bc7_blocks(rng, nb, mode)
        got, bpp = _decode(blocks, w, h, 7, 0)
        want, _ = _decode(blocks, w, h, 7, 0, ref)
        assert np.array_equal(got, want), (mode, w, h)
        assert (got[:, w * bpp:] == 0xA5).all()


@pytest.mark.parametrize("mode", BC6_MODES + BC6_RESERVED)
@pytest.mark.parametrize("signed", [0, 1])
def test_bc6h_matches_reference(mode, signed):
    ref = reference_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (no reference sources)")
    rng = np.random.default_rng(600 + mode + 100 * signed)
    flag = 0 if signed else 1  # no_alpha_or_unsigned: 1 selects the unsigned form
    for w, h in ((4, 4), (9, 6), (32, 16)):
        nb = ((w + 3) // 4) * ((h + 3) // 4)
        blocks = _bc6_blocks(rng, nb, mode)
        got, bpp = _decode(blocks, w, h, 6, flag)
        want, _ = _decode(blocks, w, h, 6, flag, ref)
        assert np.array_equal(got, want), (mode, signed, w, h)
        assert (got[:, w * bpp:] == 0xA5).all()


def _bits(fields, total=128):
    """Block bytes from (value, bit count) fields in stream order (LSB first)."""
    v, pos = 0, 0
    for val, n in fields:
        v |= (val & ((1 << n) - 1)) << pos
        pos += n
    assert pos <= total
    return np.frombuffer(v.to_bytes(16, "little"), np.uint8)


def test_bc7_mode6_equal_endpoints_is_solid():
    rng = np.random.default_rng(5)
    r, g, b, a, p = 90, 17, 120, 33, 1
    fields = [(1 << 6, 7), (r, 7), (r, 7), (g, 7), (g, 7), (b, 7), (b, 7), (a, 7), (a, 7), (p, 1), (p, 1),
              (int(rng.integers(0, 1 << 63)), 63)]
    out, _ = _decode(_bits(fields), 4, 4, 7, 0)
    px = out[:, :16].reshape(4, 4, 4)
    want = [(c << 1 | p) for c in (b, g, r, a)]
    assert (px == np.array(want, np.uint8)).all()


def test_bc7_mode6_weights():
    idx = list(range(16))
    # texel 0 is the anchor (3 bits): it takes index 0
    fields = [(1 << 6, 7)] + [(0, 7), (127, 7)] * 4 + [(0, 1), (1, 1)] + [(0, 3)] + [(i, 4) for i in idx[1:]]
    out, _ = _decode(_bits(fields), 4, 4, 7, 0)
    px = out[:, :16].reshape(16, 4)
    w4 = [0, 4, 9, 13, 17, 21, 26, 30, 34, 38, 43, 47, 51, 55, 60, 64]
    for t in range(16):
        assert (px[t] == ((64 - w4[t]) * 0 + w4[t] * 255 + 32) >> 6).all(), t


def test_bc6h_mode3_spans_the_half_range():
    fields = [(3, 5)] + [(0, 10)] * 3 + [(1023, 10)] * 3 + [(0, 3)] + [(i, 4) for i in range(1, 16)]
    out, _ = _decode(_bits(fields), 4, 4, 6, 1)
    px = out[:, :32].copy().view(np.uint16).reshape(16, 4)
    w4 = [0, 4, 9, 13, 17, 21, 26, 30, 34, 38, 43, 47, 51, 55, 60, 64]
    for t in range(16):
        v = ((0xFFFF * w4[t] + 32) >> 6) * 31 >> 6
        assert (px[t, :3] == v).all() and px[t, 3] == 0x3C00
    assert px[15, 0] == 0x7BFF


def test_invalid_blocks():
    z = np.zeros(16, np.uint8)
    out, _ = _decode(z, 4, 4, 7, 0)
    assert (out[:, :16] == 0).all()
    bad = z.copy()
    bad[0] = 19  # reserved BC6H mode
    out, _ = _decode(bad, 4, 4, 6, 1)
    px = out[:, :32].copy().view(np.uint16).reshape(16, 4)
    assert (px[:, :3] == 0).all() and (px[:, 3] == 0x3C00).all()
