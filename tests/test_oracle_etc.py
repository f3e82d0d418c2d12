"""The oracle's ETC2 / EAC decoder (oracle/swr_oracle_etc.c).

* bit for bit against the reference's ETC_Decoder.cpp compiled unmodified
  into oracle/_ref (when the reference sources are present): random blocks
  of every input type -- random bytes reach the individual, differential,
  T, H and planar modes and every punch-through and EAC table -- on ragged
  image sizes, with destination bytes outside the image left untouched;
* closed forms from the format definition: an individual-mode block whose
  indices select one modifier gives base + modifier everywhere; a planar
  block reproduces its origin at (0, 0) and its H and V colours at (4, 0)
  and (0, 4) by extrapolation of the three-point plane; punch-through
  index 2 is transparent black; EAC with multiplier 0 is base * 8 + 4 +
  modifier, shifted into 16 bits.
"""
import numpy as np
import pytest

from oracle_runner import ETC_TYPES, etc_block_bytes, etc_out_bpp, oracle_decode_etc, reference_lib


def _decode(blocks, w, h, t, lib=None, fill=0xA5):
    bpp = etc_out_bpp(t)
    pitch = w * bpp + 12
    dst = np.full((h, pitch), fill, np.uint8)
    assert oracle_decode_etc(blocks, dst, w, h, pitch, bpp, t, lib) == 1
    return dst


@pytest.mark.parametrize("name", list(ETC_TYPES))
def test_matches_reference_decoder(name):
    ref = reference_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (no reference sources)")
    t = ETC_TYPES[name]
    rng = np.random.default_rng(100 + t)
    for w, h in ((4, 4), (13, 7), (64, 32), (1, 1), (30, 5)):
        nb = ((w + 3) // 4) * ((h + 3) // 4)
        blocks = rng.integers(0, 256, nb * etc_block_bytes(t), dtype=np.uint8)
        got = _decode(blocks, w, h, t)
        want = _decode(blocks, w, h, t, ref)
        assert np.array_equal(got, want), (name, w, h)
        assert (got[:, w * etc_out_bpp(t):] == 0xA5).all()


def _word(bits):
    """8 bytes of a block from (value, hi, lo) fields of the big-endian word."""
    w = 0
    for v, hi, lo in bits:
        w |= (v & ((1 << (hi - lo + 1)) - 1)) << lo
    return np.frombuffer(w.to_bytes(8, "big"), np.uint8)


def test_individual_mode_single_modifier():
    # R1 = 9, G1 = 3, B1 = 12 (x 17 in 8 bits), table 4 (18, 60), every
    # pixel index 1 (+60): left subblock only (no flip, columns 0-1).
    blk = _word([(9, 63, 60), (3, 55, 52), (12, 47, 44), (4, 39, 37), (0, 33, 33), (0, 32, 32),
                 (0x0000, 31, 16), (0xFFFF, 15, 0)])
    out = _decode(blk, 4, 4, ETC_TYPES["rgb8"])[:, :16].reshape(4, 4, 4)
    want = np.array([min(12 * 17 + 60, 255), 3 * 17 + 60, 9 * 17 + 60, 255], np.uint8)  # B, G, R, A
    assert (out[:, :2] == want).all()


def test_planar_block_plane():
    # Blue overflows (B + dB > 31) with red and green in range: planar.
    O, Hc, V = (10, 40, 20), (50, 100, 30), (0, 127, 63)
    f = [(O[0], 62, 57), (O[1] >> 6, 56, 56), (O[1] & 63, 54, 49), (O[2] >> 5, 48, 48), ((O[2] >> 3) & 3, 44, 43),
         ((O[2] >> 1) & 3, 41, 40), (O[2] & 1, 39, 39), (Hc[0] >> 1, 38, 34), (Hc[0] & 1, 32, 32), (1, 33, 33),
         (Hc[1], 31, 25), (Hc[2], 24, 19), (V[0], 18, 13), (V[1], 12, 6), (V[2], 5, 0)]
    w0 = 0
    for v, hi, lo in f:
        w0 |= v << lo

    def sx3(v): return (v ^ 4) - 4

    def channel(w, hi):  # 5-bit base + 3-bit delta of the differential reading
        return ((w >> (hi - 4)) & 31) + sx3((w >> (hi - 7)) & 7)
    # The bits planar leaves unused (63, 55, 47-45, 42) select the mode: red
    # and green in range, blue out of it.
    for u in range(64):
        w = w0 | (u & 1) << 63 | (u >> 1 & 1) << 55 | (u >> 2 & 7) << 45 | (u >> 5 & 1) << 42
        if 0 <= channel(w, 63) <= 31 and 0 <= channel(w, 55) <= 31 and not 0 <= channel(w, 47) <= 31:
            break
    else:
        pytest.fail("no planar encoding")
    blk = np.frombuffer(w.to_bytes(8, "big"), np.uint8)
    out = _decode(blk, 4, 4, ETC_TYPES["rgb8"])[:, :16].reshape(4, 4, 4).astype(int)

    def e6(x): return (x << 2) | (x >> 4)

    def e7(x): return (x << 1) | (x >> 6)
    o = (e6(O[0]), e7(O[1]), e6(O[2]))
    hh = (e6(Hc[0]), e7(Hc[1]), e6(Hc[2]))
    vv = (e6(V[0]), e7(V[1]), e6(V[2]))
    for y in range(4):
        for x in range(4):
            for k, ch in ((0, 2), (1, 1), (2, 0)):
                want = (x * (hh[k] - o[k]) + y * (vv[k] - o[k]) + 4 * o[k] + 2) >> 2
                assert out[y, x, ch] == min(max(want, 0), 255)
    assert out[0, 0, 2] == o[0] and out[0, 0, 1] == o[1] and out[0, 0, 0] == o[2]


def test_punchthrough_index2_is_transparent():
    # Differential (opaque bit 0), indices: pixel (x, y) gets 2 where x == y.
    msb = lsb = 0
    for x in range(4):
        for y in range(4):
            k = 4 * x + y
            idx = 2 if x == y else 1
            msb |= (idx >> 1) << k
            lsb |= (idx & 1) << k
    blk = _word([(16, 63, 59), (1, 58, 56), (16, 55, 51), (0, 50, 48), (16, 47, 43), (0, 42, 40), (3, 39, 37),
                 (3, 36, 34), (0, 33, 33), (0, 32, 32), (msb, 31, 16), (lsb, 15, 0)])
    out = _decode(blk, 4, 4, ETC_TYPES["rgb8a1"])[:, :16].reshape(4, 4, 4)
    for x in range(4):
        for y in range(4):
            px = out[y, x]
            if x == y:
                assert (px == 0).all()
            else:
                r = 16 if x < 2 else 17  # no flip: columns 2-3 take R + dR
                assert px[3] == 255 and px[2] == min((r << 3 | r >> 2) + 42, 255)


def test_eac_multiplier_zero():
    table, base = 13, 200
    idx = np.arange(16) % 8
    w = base << 56 | 0 << 52 | table << 48
    for k in range(16):
        w |= int(idx[k]) << (45 - 3 * k)
    blk = np.frombuffer(w.to_bytes(8, "big"), np.uint8)
    out = _decode(blk, 4, 4, ETC_TYPES["r11"])[:, :8].copy().view(np.uint16).reshape(4, 4)
    mods = [-1, -2, -3, -10, 0, 1, 2, 9]
    for x in range(4):
        for y in range(4):
            v = base * 8 + 4 + mods[idx[4 * x + y]]
            assert out[y, x] == min(max(v, 0), 2047) << 5
