"""The oracle's BC1-BC5 decoder (BC_Decoder.cpp:31-217, 1662-1777).

Pinned bit for bit against the reference's own BC_Decoder.cpp compiled
unmodified into oracle/_ref (random blocks, every n and alpha / sign mode,
ragged image sizes, pixel strides that leave bytes untouched), and against
closed forms: 565 endpoints widen by bit replication, solid and
3-colour blocks, BC2's nibble replication, BC4's interpolation and its
signed extremes.
"""
import numpy as np
import pytest

from oracle_runner import bc_block_bytes, oracle_decode_bc, reference_lib

SHAPES = [(16, 12), (13, 7), (5, 3), (1, 1), (4, 9)]


def _modes():
    for n in range(1, 6):
        for flag in (0, 1):
            for bpp in ((4,) if n <= 3 else ((1, 2, 3) if n == 4 else (2, 4))):
                yield n, flag, bpp


def _decode(src, w, h, bpp, n, flag, lib=None, pad=3):
    pitch = w * bpp + pad
    dst = np.full(pitch * h, 0x5A, np.uint8)
    assert oracle_decode_bc(src, dst, w, h, pitch, bpp, n, flag, lib) == 1
    return dst


@pytest.mark.parametrize("n,flag,bpp", list(_modes()))
def test_matches_the_compiled_reference(n, flag, bpp):
    ref = reference_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    rng = np.random.default_rng(n * 10 + flag + 100 * bpp)
    for w, h in SHAPES:
        nb = ((w + 3) // 4) * ((h + 3) // 4)
        src = rng.integers(0, 256, nb * bc_block_bytes(n), dtype=np.uint8)
        a = _decode(src, w, h, bpp, n, flag)
        b = _decode(src, w, h, bpp, n, flag, lib=ref)
        assert np.array_equal(a, b), (w, h)


def _bc1(c0, c1, idx, flag=0):
    blk = np.array([c0 & 0xFF, c0 >> 8, c1 & 0xFF, c1 >> 8, idx & 0xFF, (idx >> 8) & 0xFF, (idx >> 16) & 0xFF,
                    idx >> 24], np.uint8)
    return _decode(blk, 4, 4, 4, 1, flag, pad=0).reshape(4, 4, 4)  # (row, col, BGRA)


def test_565_endpoints_widen_by_bit_replication():
    # Replicating the top bits is floor(v * 33 / 4) for 5 bits and
    # floor(v * 65 / 16) for 6: exact at 0 and full scale, strictly
    # increasing, within one step of v * 255 / (2^bits - 1).
    blue = np.array([_bc1(v, 0, 0)[0, 0, 0] for v in range(32)], int)  # index 0 selects c0
    red = np.array([_bc1(v << 11, 0, 0)[0, 0, 2] for v in range(32)], int)
    green = np.array([_bc1(v << 5, 0, 0)[0, 0, 1] for v in range(64)], int)
    v5, v6 = np.arange(32), np.arange(64)
    assert np.array_equal(blue, (v5 * 33) // 4) and np.array_equal(red, blue)
    assert np.array_equal(green, (v6 * 65) // 16)
    for arr, m in ((blue, 31), (green, 63)):
        assert arr[0] == 0 and arr[-1] == 255 and np.all(np.diff(arr) > 0)
        assert np.abs(arr - np.arange(m + 1) * 255 / m).max() < 1


def test_bc1_modes():
    white = _bc1(0xFFFF, 0xFFFF, 0)
    assert (white == 255).all()
    # c0 > c1: four colours, thirds of the way from c0 to c1.
    px = _bc1(0xF800, 0x001F, 0b11100100)[0]  # indices 0, 1, 2, 3 along the first row
    assert list(px[0]) == [0, 0, 255, 255] and list(px[1]) == [255, 0, 0, 255]
    assert list(px[2]) == [85, 0, 170, 255] and list(px[3]) == [170, 0, 85, 255]
    # c0 <= c1: three colours and black; with alpha, black is transparent.
    px = _bc1(0x001F, 0xF800, 0b11100100)[0]
    assert list(px[2]) == [127, 0, 127, 255] and list(px[3]) == [0, 0, 0, 0]
    px = _bc1(0x001F, 0xF800, 0b11100100, flag=1)[0]  # BC1 RGB: opaque black
    assert list(px[3]) == [0, 0, 0, 255]


def test_bc2_alpha_nibbles_replicate():
    alpha = np.arange(16, dtype=np.uint64)
    word = int(np.sum(alpha << (4 * np.arange(16, dtype=np.uint64))))
    blk = np.concatenate([np.frombuffer(np.uint64(word).tobytes(), np.uint8),
                          np.array([0xFF, 0xFF, 0xFF, 0xFF, 0, 0, 0, 0], np.uint8)])
    out = _decode(blk, 4, 4, 4, 2, 0, pad=0).reshape(16, 4)
    assert list(out[:, 3]) == [a * 17 for a in range(16)]


def _bc4(c0, c1, codes, signed):
    word = c0 | (c1 << 8)
    for t, c in enumerate(codes):
        word |= c << (16 + 3 * t)
    blk = np.frombuffer(np.uint64(word).tobytes(), np.uint8)
    return _decode(blk, 4, 4, 1, 4, 0 if signed else 1, pad=0)


def test_bc4_interpolation_and_extremes():
    codes = list(range(8)) * 2
    out = _bc4(200, 60, codes, signed=False)[:8].astype(int)
    assert list(out) == [200, 60] + [((8 - i) * 200 + (i - 1) * 60) // 7 for i in range(2, 8)]
    assert np.all(np.diff(out[[0, 2, 3, 4, 5, 6, 7, 1]]) < 0)  # c0 down to c1
    out = _bc4(60, 200, codes, signed=False)[:8].astype(int)
    assert list(out[6:]) == [0, 255]
    out = _bc4(0xF0, 0x10, codes, signed=True)[:8].astype(np.int8)  # -16 <= 16: six codes plus extremes
    assert list(out[6:]) == [-128, 127]
    assert list(out[2:6]) == [int((4 * -16 + 1 * 16) / 5), int((3 * -16 + 2 * 16) / 5), int((2 * -16 + 3 * 16) / 5),
                              int((1 * -16 + 4 * 16) / 5)]  # truncation toward zero
