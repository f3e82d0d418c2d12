"""The oracle's mirrored address modes (SamplerCore.cpp:2398-2435 for the
16-bit path, 2480-2630 for the float path; SpirvShaderSampling.cpp:419-430)
against the textbook wrap functions.

A screen-wide quad sweeps u from -2 to 2 across four periods of an 8x1
texture, one pixel per texel centre, with nearest filtering and one mip:
pixel x of period n (-2..1) and texel i = x mod 8 must show
  REPEAT                texel i;
  MIRRORED_REPEAT       texel i for even n, 7 - i for odd n;
  MIRROR_CLAMP_TO_EDGE  7 - i for n = -1, i for n = 0, the edge texel 7
                        beyond |u| = 1,
for an RGBA8 texture (16-bit path) and an RGBA32F texture (float path).
"""
import numpy as np
import pytest

from oracle_runner import oracle_render
from swiftshader_amd import abi
from swiftshader_amd.scenes import DrawData, TextureData

TW = 8


def _render(mode, float_tex):
    tex8 = (np.arange(TW, dtype=np.uint8)[:, None] * np.array([30, 7, 1, 0], np.uint8) + np.array([5, 9, 200, 255],
                                                                                                np.uint8))
    tex8 = tex8[None]  # (1, 8, 4)
    level = tex8.astype(np.float32) / 255.0 if float_tex else tex8
    tex = TextureData(format=abi.FORMAT_R32G32B32A32_SFLOAT if float_tex else abi.FORMAT_R8G8B8A8_UNORM,
                      levels=[level], mag_filter=abi.FILTER_NEAREST, min_filter=abi.FILTER_NEAREST,
                      mipmap_mode=abi.MIPMAP_MODE_NEAREST, address_u=mode, address_v=abi.ADDRESS_CLAMP_TO_EDGE)
    xy = np.array([[-1, -1], [1, -1], [1, 1], [-1, -1], [1, 1], [-1, 1]], np.float64)
    pos = np.concatenate([xy, np.full((6, 1), 0.5), np.ones((6, 1))], axis=1).astype(np.float32)
    uv = np.stack([xy[:, 0] * 2.0, np.full(6, 0.5)], 1).astype(np.float32)  # u in [-2, 2]
    d = DrawData(width=4 * TW, height=1, positions=pos, colors=np.ones((6, 4), np.float32), texcoords=uv,
                 shader=abi.SHADER_SAMPLE, texture=tex, color_format=abi.FORMAT_R32G32B32A32_SFLOAT)
    got = oracle_render(d)[0][0, 0]  # (4 TW, 4)
    return got, level[0]


def _want_texel(mode, x):
    n, i = x // TW - 2, x % TW
    if mode == abi.ADDRESS_REPEAT:
        return i
    if mode == abi.ADDRESS_MIRRORED_REPEAT:
        return i if n % 2 == 0 else TW - 1 - i
    return {-2: TW - 1, -1: TW - 1 - i, 0: i, 1: TW - 1}[n]  # MIRROR_CLAMP_TO_EDGE


@pytest.mark.parametrize("float_tex", [False, True], ids=["rgba8", "rgba32f"])
@pytest.mark.parametrize("mode", [abi.ADDRESS_REPEAT, abi.ADDRESS_MIRRORED_REPEAT, abi.ADDRESS_MIRROR_CLAMP_TO_EDGE],
                         ids=["repeat", "mirrored_repeat", "mirror_clamp"])
def test_nearest_texel_follows_the_wrap_function(mode, float_tex):
    got, level = _render(mode, float_tex)
    scale = 1.0 if float_tex else 255.0
    for x in range(4 * TW):
        want = level[_want_texel(mode, x)].astype(np.float64) / (1.0 if float_tex else 255.0)
        assert np.allclose(got[x], want, atol=1e-6 if float_tex else 1.0 / 255 / 2), (mode, x, got[x] * scale)
