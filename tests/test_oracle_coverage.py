"""Pins the oracle's coverage (edge DDA, spans, quad masks, MSAA masks).

Independent characterisation: for a triangle with exact 28.4 vertices, a
sample point (px, py) is covered iff the horizontal line y = py crosses the
triangle's boundary -- each edge taken half-open in y, [Ylo, Yhi) -- at two
points xa <= xb with xa <= px < xb.  That is the scanline rule with the
left-inclusive / right-exclusive and top-inclusive / bottom-exclusive ties
the reference's DDA implements (SetupRoutine.cpp:550-621), computed here
with exact rationals and sorted crossings instead of oriented edges.

Vertices are chosen on the 28.4 grid and mapped to clip space so that the
projection reproduces them exactly (viewport 64 wide: X = 512 x + 504).
"""
from fractions import Fraction

import numpy as np

from oracle_runner import oracle_render
from swiftshader_amd import abi
from swiftshader_amd.scenes import DrawData

SIZE = 64
XF = [-2, 6, -6, 2]
YF = [-6, -2, 2, 6]


def to_clip(X, Y):
    # X0xF = 32*16 - 8 = 504, WxF = 32*16 = 512 for a 64-pixel viewport.
    return np.array([(X - 504) / 512.0, (Y - 504) / 512.0, 0.5, 1.0], np.float32)


def covered(verts, px, py):
    xs = []
    for i in range(3):
        (Xa, Ya), (Xb, Yb) = verts[i], verts[(i + 1) % 3]
        if Ya == Yb:
            continue
        (Xl, Yl), (Xh, Yh) = ((Xa, Ya), (Xb, Yb)) if Ya < Yb else ((Xb, Yb), (Xa, Ya))
        if Yl <= py < Yh:
            xs.append(Fraction(Xl) + Fraction((py - Yl) * (Xh - Xl), Yh - Yl))
    if len(xs) != 2:
        return False
    xa, xb = sorted(xs)
    return xa <= px < xb


def render_single(verts, samples=1):
    pos = np.stack([to_clip(X, Y) for X, Y in verts])
    col = np.ones((3, 4), np.float32)
    d = DrawData(width=SIZE, height=SIZE, positions=pos, colors=col, samples=samples,
                 shader=abi.SHADER_INTERPOLATE, depth_test=0, depth_write=0, has_depth=False)
    c, _, _, occ = oracle_render(d)
    return c[..., 0] == 255, occ


def random_grid_triangles(seed, n, lo=0, hi=SIZE * 16):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        pts = rng.integers(lo, hi, (3, 2))
        # Snap some coordinates onto pixel centres and sample rows to hit ties.
        snap = rng.random((3, 2)) < 0.3
        pts = np.where(snap, (pts // 16) * 16, pts)
        yield [(int(a), int(b)) for a, b in pts]


def test_single_sample_coverage_matches_scanline_rule():
    for verts in random_grid_triangles(3, 150):
        got, occ = render_single(verts)
        want = np.zeros((SIZE, SIZE), bool)
        ys = [v[1] for v in verts]
        xs = [v[0] for v in verts]
        for y in range(max(0, min(ys) // 16 - 1), min(SIZE, max(ys) // 16 + 2)):
            for x in range(max(0, min(xs) // 16 - 1), min(SIZE, max(xs) // 16 + 2)):
                want[y, x] = covered(verts, 16 * x, 16 * y)
        assert np.array_equal(got[0], want), verts
        assert occ == int(want.sum())


def test_msaa_sample_masks_match_scanline_rule():
    for verts in random_grid_triangles(4, 60, lo=64, hi=SIZE * 16 - 64):
        got, occ = render_single(verts, samples=4)
        total = 0
        ys = [v[1] for v in verts]
        xs = [v[0] for v in verts]
        for q in range(4):
            want = np.zeros((SIZE, SIZE), bool)
            for y in range(max(0, min(ys) // 16 - 1), min(SIZE, max(ys) // 16 + 2)):
                for x in range(max(0, min(xs) // 16 - 1), min(SIZE, max(xs) // 16 + 2)):
                    want[y, x] = covered(verts, 16 * x + XF[q], 16 * y + YF[q])
            assert np.array_equal(got[q], want), (verts, q)
            total += int(want.sum())
        assert occ == total


def test_shared_edges_are_watertight():
    # A jittered, tessellated full-viewport mesh: every pixel is covered by
    # exactly one triangle (stencil INCREMENT on every covered sample).
    from cases import _indexed
    d = _indexed()
    d.has_stencil = True
    d.stencil_test = 1
    d.front = d.back = dict(compare_op=abi.COMPARE_ALWAYS, pass_op=abi.STENCIL_OP_INCREMENT_AND_WRAP,
                            fail_op=abi.STENCIL_OP_KEEP, depth_fail_op=abi.STENCIL_OP_INCREMENT_AND_WRAP)
    d.clear_stencil = 0
    _, _, s, occ = oracle_render(d)
    assert (s == 1).all()
    assert occ == d.width * d.height


def test_strip_and_fan_cover_their_area_once():
    """A jittered multi-row strip (with degenerate joins) and a fan around an
    interior hub tile their area: no sample is covered twice (stencil
    INCREMENT) and, away from the outline, none is missed -- which fails if
    the strip's alternating winding or the fan's hub order is wrong."""
    from cases import _fan, _strip_band
    for d in (_strip_band(seed=40), _fan(seed=41, provoking=abi.PROVOKING_VERTEX_FIRST),
              _fan(seed=42, provoking=abi.PROVOKING_VERTEX_LAST)):
        d.has_stencil = True
        d.stencil_test = 1
        d.depth_test = 0
        d.has_depth = False
        d.front = d.back = dict(compare_op=abi.COMPARE_ALWAYS, pass_op=abi.STENCIL_OP_INCREMENT_AND_WRAP,
                                fail_op=abi.STENCIL_OP_KEEP, depth_fail_op=abi.STENCIL_OP_INCREMENT_AND_WRAP)
        d.clear_stencil = 0
        _, _, s, occ = oracle_render(d)
        assert s.max() == 1
        assert occ == int((s == 1).sum()) > 500
    # The strip spans the whole target: every pixel is covered exactly once.
    d = _strip_band(seed=43)
    d.has_stencil, d.stencil_test, d.depth_test, d.has_depth = True, 1, 0, False
    d.front = d.back = dict(compare_op=abi.COMPARE_ALWAYS, pass_op=abi.STENCIL_OP_INCREMENT_AND_WRAP,
                            fail_op=abi.STENCIL_OP_KEEP, depth_fail_op=abi.STENCIL_OP_INCREMENT_AND_WRAP)
    _, _, s, _ = oracle_render(d)
    assert (s == 1).all()


def test_provoking_last_is_the_rotated_list():
    """A triangle list in last-vertex mode renders as the first-vertex list
    with each triangle's vertices rotated to (v2, v0, v1) (Renderer.cpp:100-111)."""
    from cases import _list_provoking_last
    d = _list_provoking_last()
    a = oracle_render(d)
    t = d.triangle_count
    order = (3 * np.arange(t)[:, None] + np.array([2, 0, 1])[None, :]).reshape(-1)
    import dataclasses
    e = dataclasses.replace(d, provoking_vertex=abi.PROVOKING_VERTEX_FIRST, positions=d.positions[order],
                            colors=d.colors[order], texcoords=d.texcoords[order])
    b = oracle_render(e)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32)) and a[3] == b[3]
