"""The oracle's advanced blend modes (VK_EXT_blend_operation_advanced;
PixelRoutine.cpp:1395-1615, 1937-1953) against the Vulkan specification's
formulas in FP64: premultiplied inputs with uncorrelated overlap,
X = Y = Z = 1, so  RGB = f(Cs, Cd) * as * ad + Cs * as * (1 - ad) + Cd * ad * (1 - as)
with Cs, Cd the un-premultiplied colours, and A = as + ad - as * ad.

One draw per case: the destination is the RGBA32F clear value, the source
a flat quad (no clamp on a float target).  Colours stay in [0, 1], where the
reference's single-branch ClipColor equals the specification's.
"""
import numpy as np
import pytest

from oracle_runner import oracle_render
from swiftshader_amd import abi
from swiftshader_amd.scenes import DrawData


def _lum(c):
    return 0.3 * c[0] + 0.59 * c[1] + 0.11 * c[2]


def _clip(c):
    l, n, x = _lum(c), min(c), max(c)
    if n < 0:
        c = l + (c - l) * l / (l - n)
    if x > 1:
        c = l + (c - l) * (1 - l) / (x - l)
    return c


def _set_lum(c, l):
    return _clip(c + (l - _lum(c)))


def _set_sat(c, s):
    n, x = min(c), max(c)
    return (c - n) * s / (x - n) if x > n else np.zeros(3)


def _f(op, cs, cd):
    if op == abi.BLEND_OP_MULTIPLY_EXT:
        return cs * cd
    if op == abi.BLEND_OP_SCREEN_EXT:
        return cs + cd - cs * cd
    if op == abi.BLEND_OP_OVERLAY_EXT:
        return np.where(cd <= 0.5, 2 * cs * cd, 1 - 2 * (1 - cs) * (1 - cd))
    if op == abi.BLEND_OP_DARKEN_EXT:
        return np.minimum(cs, cd)
    if op == abi.BLEND_OP_LIGHTEN_EXT:
        return np.maximum(cs, cd)
    if op == abi.BLEND_OP_COLORDODGE_EXT:
        with np.errstate(divide="ignore"):
            return np.where(cd <= 0, 0.0, np.where(cs < 1, np.minimum(1.0, cd / np.where(cs < 1, 1 - cs, 1)), 1.0))
    if op == abi.BLEND_OP_COLORBURN_EXT:
        with np.errstate(divide="ignore"):
            return np.where(cd >= 1, 1.0, np.where(cs > 0, 1 - np.minimum(1.0, (1 - cd) / np.where(cs > 0, cs, 1)), 0.0))
    if op == abi.BLEND_OP_HARDLIGHT_EXT:
        return np.where(cs <= 0.5, 2 * cs * cd, 1 - 2 * (1 - cs) * (1 - cd))
    if op == abi.BLEND_OP_SOFTLIGHT_EXT:
        return np.where(cs <= 0.5, cd - (1 - 2 * cs) * cd * (1 - cd),
                        np.where(cd <= 0.25, cd + (2 * cs - 1) * cd * ((16 * cd - 12) * cd + 3),
                                 cd + (2 * cs - 1) * (np.sqrt(cd) - cd)))
    if op == abi.BLEND_OP_DIFFERENCE_EXT:
        return np.abs(cs - cd)
    if op == abi.BLEND_OP_EXCLUSION_EXT:
        return cs + cd - 2 * cs * cd
    if op == abi.BLEND_OP_HSL_HUE_EXT:
        return _set_lum(_set_sat(cs, max(cd) - min(cd)), _lum(cd))
    if op == abi.BLEND_OP_HSL_SATURATION_EXT:
        return _set_lum(_set_sat(cd, max(cs) - min(cs)), _lum(cd))
    if op == abi.BLEND_OP_HSL_COLOR_EXT:
        return _set_lum(cs, _lum(cd))
    return _set_lum(cd, _lum(cs))  # HSL_LUMINOSITY


def _spec(op, src, dst):
    sa, da = src[3], dst[3]
    cs = src[:3] / sa if sa != 0 else np.zeros(3)
    cd = dst[:3] / da if da != 0 else np.zeros(3)
    rgb = _f(op, cs, cd) * sa * da + cs * sa * (1 - da) + cd * da * (1 - sa)
    return np.append(rgb, sa + da - sa * da)


def _render(op, src, dst):
    xy = np.array([[-1, -1], [1, -1], [1, 1], [-1, -1], [1, 1], [-1, 1]], np.float64)
    pos = np.concatenate([xy, np.full((6, 1), 0.5), np.ones((6, 1))], axis=1).astype(np.float32)
    d = DrawData(width=4, height=4, positions=pos, colors=np.repeat(np.asarray(src, np.float32)[None], 6, 0),
                 color_format=abi.FORMAT_R32G32B32A32_SFLOAT, clear_color=tuple(float(x) for x in dst),
                 blend_enable=1, blend=(abi.BLEND_FACTOR_ZERO, abi.BLEND_FACTOR_ZERO, op, abi.BLEND_FACTOR_ZERO,
                                        abi.BLEND_FACTOR_ZERO, op))
    return oracle_render(d)[0][0, 0, 0].astype(np.float64)


@pytest.mark.parametrize("op", abi.ADVANCED_BLEND_OPS, ids=lambda o: str(o - 1000148000))
def test_modes_match_the_specification(op):
    rng = np.random.default_rng(op)
    for _ in range(60):
        # premultiplied inputs: colour <= alpha, so Cs, Cd lie in [0, 1]
        sa, da = rng.uniform(0.15, 1.0, 2)
        src = np.append(rng.uniform(0, 1, 3) * sa, sa).astype(np.float32).astype(np.float64)
        dst = np.append(rng.uniform(0, 1, 3) * da, da).astype(np.float32).astype(np.float64)
        got = _render(op, src, dst)
        want = _spec(op, src, dst)
        assert np.abs(got - want).max() < 2e-5, (op, src, dst, got, want)


def test_zero_alpha_colours_count_as_zero():
    op = abi.BLEND_OP_SCREEN_EXT
    got = _render(op, np.array([0.5, 0.2, 0.9, 0.0]), np.array([0.3, 0.6, 0.1, 0.5]))
    assert np.allclose(got, _spec(op, np.array([0.5, 0.2, 0.9, 0.0]), np.array([0.3, 0.6, 0.1, 0.5])), atol=1e-6)


def test_mixed_advanced_and_basic_ops_are_refused():
    xy = np.zeros((3, 4), np.float32)
    xy[:, 3] = 1
    d = DrawData(width=4, height=4, positions=xy, colors=np.ones((3, 4), np.float32), blend_enable=1,
                 blend=(abi.BLEND_FACTOR_ONE, abi.BLEND_FACTOR_ONE, abi.BLEND_OP_MULTIPLY_EXT, abi.BLEND_FACTOR_ONE,
                        abi.BLEND_FACTOR_ONE, abi.BLEND_OP_ADD))
    with pytest.raises(RuntimeError):
        oracle_render(d)
