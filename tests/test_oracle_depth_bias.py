"""The oracle's depth bias (SetupRoutine.cpp:417-475, PixelRoutine.cpp:141-144,
Context.cpp:640-661) against closed forms.

A full-viewport triangle pair at one depth has a zero slope and, with the
depth range [0, 1], differences z1 - z0 = z2 - z0 = 0, so the constant term
is exactly r * factor with r = 2^(e - 23) from the depth's own exponent.
The slope term is checked on a tilted plane against max(|dz/dx|, |dz/dy|)
from the vertices in FP64.  The clamp, its sign rule, a NaN clamp and the
disabled state are checked exactly.  The exponent of the two other terms,
which the reference forms from z1 - z0 and z2 - z0 rather than z1 and z2, is
restated from the source and compared with the GPU only (parity unpinned
beyond that).
"""
import numpy as np

from oracle_runner import oracle_render
from swiftshader_amd import abi
from swiftshader_amd.scenes import DrawData


def _quad(z, w=1.0):
    """Two triangles covering the viewport at depth z (clip space, w)."""
    xy = np.array([[-1, -1], [1, -1], [1, 1], [-1, -1], [1, 1], [-1, 1]], np.float64)
    pos = np.concatenate([xy * w, np.full((6, 1), z * w), np.full((6, 1), w)], axis=1).astype(np.float32)
    col = np.ones((6, 4), np.float32)
    return pos, col


def _depth(pos, col, bias, size=16, compare=abi.COMPARE_ALWAYS):
    d = DrawData(width=size, height=size, positions=pos, colors=col, depth_compare=compare, depth_bias=bias)
    _, depth, _, _ = oracle_render(d)
    return depth[0]


def test_constant_term_is_factor_times_the_depth_ulp():
    for z, k in ((0.5, 3.0), (0.3, -7.0), (0.75, 1000.0), (0.0625, 1.0)):
        pos, col = _quad(z)
        got = _depth(pos, col, (k, 0.0, 0.0))
        zf = np.float32(z)
        e = int(np.floor(np.log2(zf)))
        want = np.float32(zf + np.float32(k) * np.float32(2.0 ** (e - 23)))
        assert (got == want).all(), (z, k, got[0, 0], want)


def test_slope_term_follows_the_plane():
    # z varies along x only: z = 0.2 + 0.5 * (x_ndc + 1) / 2 over 32 px.
    size = 32
    xy = np.array([[-1, -1], [1, -1], [1, 1], [-1, -1], [1, 1], [-1, 1]], np.float64)
    z = 0.2 + 0.25 * (xy[:, 0] + 1.0)
    pos = np.concatenate([xy, z[:, None], np.ones((6, 1))], axis=1).astype(np.float32)
    col = np.ones((6, 4), np.float32)
    base = _depth(pos, col, None, size)
    biased = _depth(pos, col, (0.0, 0.0, 4.0), size)
    slope = 0.5 / size  # dz per pixel
    assert np.allclose(biased.astype(np.float64) - base, 4.0 * slope, rtol=0, atol=2e-7)


def test_clamp_rules():
    pos, col = _quad(0.5)
    # positive clamp caps a large positive bias
    got = _depth(pos, col, (1e6, 0.004, 0.0))
    assert (got == np.float32(np.float32(0.5) + np.float32(0.004))).all()
    # negative clamp floors a large negative bias
    got = _depth(pos, col, (-1e7, -0.125, 0.0))
    assert (got == np.float32(0.375)).all()
    # a clamp of the other sign does not bind: min(bias, c) with c > 0 and bias < 0
    got = _depth(pos, col, (-4.0, 0.5, 0.0))
    assert (got == np.float32(np.float32(0.5) - np.float32(4.0 * 2.0 ** -24))).all()
    # a NaN clamp is no clamp
    got = _depth(pos, col, (8.0, float("nan"), 0.0))
    assert (got == np.float32(np.float32(0.5) + np.float32(8.0 * 2.0 ** -24))).all()


def test_disabled_or_zero_factors_change_nothing():
    rng = np.random.default_rng(5)
    pos = np.concatenate([rng.uniform(-1, 1, (60, 2)), rng.uniform(0, 1, (60, 1)), np.ones((60, 1))],
                         axis=1).astype(np.float32)
    col = rng.uniform(0, 1, (60, 4)).astype(np.float32)
    base = _depth(pos, col, None, 24, abi.COMPARE_LESS)
    assert np.array_equal(base.view(np.uint32), _depth(pos, col, (0.0, 0.3, 0.0), 24, abi.COMPARE_LESS).view(np.uint32))


def test_bias_decides_visibility_between_coplanar_draws():
    # The same surface drawn twice: with LESS the second draw loses at equal
    # depth, and wins once its depth is biased towards the viewer.
    pos, col = _quad(0.5)
    d1 = DrawData(width=8, height=8, positions=pos, colors=col, clear_color=(0, 0, 0, 0))
    color, depth, stencil, _ = oracle_render(d1)
    red = np.tile(np.array([1, 0, 0, 1], np.float32), (6, 1))
    d2 = DrawData(width=8, height=8, positions=pos, colors=red)
    c2, _, _, occ = oracle_render(d2, targets=(color.copy(), depth.copy(), stencil))
    assert occ == 0 and (c2[..., 0] == 255).all()  # white stays (first draw's colour is white)
    d3 = DrawData(width=8, height=8, positions=pos, colors=red, depth_bias=(-2.0, 0.0, 0.0))
    c3, _, _, occ = oracle_render(d3, targets=(color.copy(), depth.copy(), stencil))
    assert occ == 64 and (c3[..., 1] == 0).all()
