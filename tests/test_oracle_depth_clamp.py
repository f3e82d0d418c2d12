"""The oracle's depth clamp and depth clip (Context.cpp:647-648, 686-693;
PixelProcessor.cpp:120-136; VertexRoutine.cpp:137-143; Renderer.cpp:940,
1176) against closed forms.

A flat quad at NDC depth 1.5 lies beyond the far plane.  With depth
clipping (the default without depth clamp) it disappears; without it the
quad is rasterised and its depth, 1.5 * range + near, is clamped to the
viewport's depth range when depth clamp is on and to [0, 1] otherwise.
Lines and points follow the same clip state.
"""
import numpy as np
import pytest

from oracle_runner import oracle_render
from swiftshader_amd import abi
from swiftshader_amd.scenes import DrawData

W, H = 12, 10
VP = (0.0, 0.0, float(W), float(H), 0.2, 0.7)


def _quad(z):
    xy = np.array([[-1, -1], [1, -1], [1, 1], [-1, -1], [1, 1], [-1, 1]], np.float64)
    return np.concatenate([xy, np.full((6, 1), z), np.ones((6, 1))], axis=1).astype(np.float32)


def _draw(pos, clamp=0, clip=0, topology=abi.TOPOLOGY_TRIANGLE_LIST, viewport=VP):
    d = DrawData(width=W, height=H, positions=pos, colors=np.ones((len(pos), 4), np.float32), viewport=viewport,
                 depth_compare=abi.COMPARE_ALWAYS, depth_clamp_enable=clamp, depth_clip_control=clip,
                 topology=topology, clear_depth=1.0)
    _, depth, _, occ = oracle_render(d)
    return depth[0], occ


def test_beyond_far_is_clipped_by_default():
    depth, occ = _draw(_quad(1.5))
    assert occ == 0 and (depth == 1.0).all()


def test_depth_clamp_keeps_the_quad_at_the_far_depth():
    depth, occ = _draw(_quad(1.5), clamp=1)
    assert occ == W * H and (depth == np.float32(0.7)).all()


def test_clip_off_without_clamp_clamps_to_unit_range():
    depth, occ = _draw(_quad(1.5), clip=2)
    want = np.float32(np.float32(1.5) * np.float32(0.5)) + np.float32(0.2)  # 0.95, inside [0, 1]
    assert occ == W * H and (depth == want).all()
    depth, occ = _draw(_quad(3.0), clip=2)  # 1.7 -> clamped to 1
    assert occ == W * H and (depth == 1.0).all()


def test_clip_on_with_clamp_clips():
    depth, occ = _draw(_quad(1.5), clamp=1, clip=1)
    assert occ == 0


def test_in_range_depth_is_unchanged_by_the_clamp():
    a, _ = _draw(_quad(0.3))
    b, _ = _draw(_quad(0.3), clamp=1)
    assert np.array_equal(a, b)


def test_reversed_viewport_range_clamps_to_its_minimum_and_maximum():
    vp = (0.0, 0.0, float(W), float(H), 0.9, 0.4)  # range -0.5: NDC 1.5 -> 0.15
    depth, occ = _draw(_quad(1.5), clamp=1, viewport=vp)
    assert occ == W * H and (depth == np.float32(0.4)).all()


@pytest.mark.parametrize("topology", [abi.TOPOLOGY_LINE_LIST, abi.TOPOLOGY_POINT_LIST])
def test_lines_and_points_follow_the_clip_state(topology):
    pts = np.array([[-0.8, -0.2, 1.5, 1.0], [0.7, 0.3, 1.5, 1.0]], np.float32)
    _, occ = _draw(pts, topology=topology)
    assert occ == 0
    depth, occ = _draw(pts, clamp=1, topology=topology)
    assert occ > 0 and set(np.unique(depth[depth < 1.0])) == {np.float32(0.7)}


def test_invalid_clip_control_is_refused():
    with pytest.raises(RuntimeError):
        _draw(_quad(0.5), clip=3)
