"""The oracle's points (DrawCall::setupPoint, Renderer.cpp:1166-1210;
SetupRoutine.cpp:52-54, 115-119, 393-409) against closed forms.

Pixel (col, row) spans [col, col + 1) x [row, row + 1), its centre at +0.5;
coverage is read from the depth attachment (cleared to 1).

- A square of integer side k covers exactly k x k pixels wherever it lies
  inside the target (k pixel centres fall in any half-open interval of
  length k), whatever the vertex's w; at 4 samples each sample plane holds
  k x k.
- A side below 1 reads as 1 and above 1023 as 1023; a NaN side is dropped.
- Depth is the vertex's z / w mapped through the depth range, exactly, and
  the colour is the vertex's colour.
- A point across the target's edge covers exactly its visible part.
"""
import numpy as np
import pytest

from oracle_runner import oracle_render
from swiftshader_amd import abi
from swiftshader_amd.scenes import DrawData

W, H = 40, 32


def _pt(px, py, w=1.0, z=0.5):
    return [(2.0 * px / W - 1.0) * w, (2.0 * py / H - 1.0) * w, z * w, w]


def _draw(points, sizes=None, colors=None, samples=1, viewport=None):
    pos = np.asarray(points, np.float32).reshape(-1, 4)
    col = np.ones((len(pos), 4), np.float32) if colors is None else np.asarray(colors, np.float32)
    d = DrawData(width=W, height=H, positions=pos, colors=col, shader=abi.SHADER_INTERPOLATE, samples=samples,
                 color_format=abi.FORMAT_R32G32B32A32_SFLOAT, topology=abi.TOPOLOGY_POINT_LIST,
                 point_sizes=None if sizes is None else np.asarray(sizes, np.float32),
                 depth_compare=abi.COMPARE_ALWAYS, viewport=viewport)
    color, depth, _, occ = oracle_render(d)
    return color, depth, occ


def test_integer_squares_cover_k_by_k_pixels():
    rng = np.random.default_rng(3)
    for _ in range(60):
        k = int(rng.integers(1, 9))
        w = float(rng.uniform(0.3, 4.0))
        px, py = rng.uniform(k / 2 + 0.5, [W - k / 2 - 0.5, H - k / 2 - 0.5])
        _, depth, occ = _draw([_pt(px, py, w)], sizes=[k])
        cov = depth[0] < 1.0
        rows, cols = np.nonzero(cov)
        assert cov.sum() == k * k and occ == k * k, (k, px, py, w)
        assert rows.max() - rows.min() == k - 1 and cols.max() - cols.min() == k - 1


def test_centred_squares_cover_the_pixels_around_the_centre():
    _, depth, _ = _draw([_pt(10.5, 7.5)], sizes=[1.0])
    want = np.zeros((H, W), bool)
    want[7, 10] = True
    assert np.array_equal(depth[0] < 1.0, want)
    _, depth, _ = _draw([_pt(20.5, 15.5)], sizes=[3.0])
    want = np.zeros((H, W), bool)
    want[14:17, 19:22] = True
    assert np.array_equal(depth[0] < 1.0, want)
    _, depth, _ = _draw([_pt(6.0, 9.0)], sizes=[2.0])  # on a pixel corner
    want = np.zeros((H, W), bool)
    want[8:10, 5:7] = True
    assert np.array_equal(depth[0] < 1.0, want)


def test_each_sample_plane_holds_k_by_k():
    rng = np.random.default_rng(4)
    for _ in range(20):
        k = int(rng.integers(1, 6))
        px, py = rng.uniform(k / 2 + 1, [W - k / 2 - 1, H - k / 2 - 1])
        _, depth, occ = _draw([_pt(px, py)], sizes=[k], samples=4)
        assert [(depth[s] < 1.0).sum() for s in range(4)] == [k * k] * 4 and occ == 4 * k * k


def test_sizes_are_clamped_and_nan_is_dropped():
    a = _draw([_pt(12.3, 9.8)], sizes=[0.25])
    b = _draw([_pt(12.3, 9.8)], sizes=[1.0])
    c = _draw([_pt(12.3, 9.8)])  # no sizes: 1
    assert np.array_equal(a[1], b[1]) and np.array_equal(b[1], c[1])
    big = _draw([_pt(20.0, 16.0)], sizes=[5000.0])
    assert (big[1][0] < 1.0).all()  # 1023 covers the target
    _, depth, occ = _draw([_pt(20.0, 16.0)], sizes=[np.nan])
    assert occ == 0 and (depth[0] == 1.0).all()


def test_depth_and_colour_are_the_vertex_values():
    rng = np.random.default_rng(5)
    pts, cols, zs = [], [], []
    for i in range(6):
        z, w = float(rng.uniform(0.05, 0.95)), float(rng.uniform(0.5, 3.0))
        pts.append([(2.0 * (4.5 + 6 * i) / W - 1.0) * w, (2.0 * (10.5 + 2 * i) / H - 1.0) * w,
                    float(np.float32(np.float32(z) * np.float32(w))), w])
        cols.append(rng.uniform(0, 1, 4))
        rhw = np.float32(1.0) / np.float32(w)
        zs.append(np.float32(np.float32(np.float32(z) * np.float32(w)) * rhw))  # z * (1 / w), as projected
    color, depth, _ = _draw(pts, sizes=[3] * 6, colors=cols, viewport=(0.0, 0.0, W, H, 0.25, 0.75))
    for i in range(6):
        r, c = 10 + 2 * i, 4 + 6 * i
        block = (slice(r - 1, r + 2), slice(c - 1, c + 2))
        want_z = np.float32(zs[i] * np.float32(0.5) + np.float32(0.25))
        assert (depth[0][block] == want_z).all(), i
        assert np.abs(color[0][block] - cols[i]).max() < 1e-6, i


def test_point_across_the_edge_covers_its_visible_part():
    _, depth, _ = _draw([_pt(1.0, 30.0)], sizes=[6.0])  # x in [-2, 4), y in [27, 33)
    want = np.zeros((H, W), bool)
    want[27:32, 0:4] = True
    assert np.array_equal(depth[0] < 1.0, want)


def test_refused_point_states():
    d = DrawData(width=W, height=H, positions=np.asarray([_pt(5, 5)], np.float32), colors=np.ones((1, 4), np.float32),
                 topology=abi.TOPOLOGY_POINT_LIST, depth_bias=(0.0, 0.0, 2.0))
    with pytest.raises(RuntimeError):
        oracle_render(d)
