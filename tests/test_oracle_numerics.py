"""Pins the oracle's depth, interpolation, blend and sampler arithmetic
against FP64 closed forms (textbook / Vulkan-spec formulas), with the
tolerances the reference's float32 arithmetic allows.

Each check is chosen so that a plausible mistake (a swapped plane
coefficient, a missing perspective divide, a wrong blend factor, an
off-by-half texel offset, a wrong mip weight) fails it.
"""
import numpy as np
import pytest

from oracle_runner import oracle_render
from swiftshader_amd import abi, scenes
from swiftshader_amd.scenes import DrawData, TextureData

SIZE = 64


def proj(pos, size=SIZE):
    """Viewport transform to pixel units (pixel centres are integers), with
    the vertices snapped to the 28.4 sub-pixel grid as the rasteriser sees them."""
    x, y, z, w = pos.astype(np.float64).T
    X = np.rint(((x / w + 1) * 0.5 * size - 0.5) * 16) / 16
    Y = np.rint(((y / w + 1) * 0.5 * size - 0.5) * 16) / 16
    return X, Y, z / w, w


def barycentrics(X, Y, px, py):
    d = (Y[1] - Y[2]) * (X[0] - X[2]) + (X[2] - X[1]) * (Y[0] - Y[2])
    l0 = ((Y[1] - Y[2]) * (px - X[2]) + (X[2] - X[1]) * (py - Y[2])) / d
    l1 = ((Y[2] - Y[0]) * (px - X[2]) + (X[0] - X[2]) * (py - Y[2])) / d
    return l0, l1, 1 - l0 - l1


def random_triangle(rng):
    while True:
        w = rng.uniform(0.5, 3.0, 3)
        xy = rng.uniform(-0.9, 0.9, (3, 2))
        z = rng.uniform(0.05, 0.95, 3)
        pos = np.stack([xy[:, 0] * w, xy[:, 1] * w, z * w, w], 1).astype(np.float32)
        X, Y, _, _ = proj(pos)
        area = 0.5 * abs((X[1] - X[0]) * (Y[2] - Y[0]) - (X[2] - X[0]) * (Y[1] - Y[0]))
        if area > 60:
            return pos


def test_depth_matches_fp64_plane():
    rng = np.random.default_rng(8)
    for _ in range(40):
        pos = random_triangle(rng)
        col = np.ones((3, 4), np.float32)
        d = DrawData(width=SIZE, height=SIZE, positions=pos, colors=col, depth_compare=abi.COMPARE_ALWAYS)
        _, z, _, occ = oracle_render(d)
        X, Y, Z, _ = proj(pos)
        ys, xs = np.nonzero(z[0] < 1.0)
        assert len(ys) == occ > 0
        l0, l1, l2 = barycentrics(X, Y, xs.astype(np.float64), ys.astype(np.float64))
        want = np.clip(l0 * Z[0] + l1 * Z[1] + l2 * Z[2], 0, 1)
        got = z[0][ys, xs].astype(np.float64)
        # Screen-linear depth; float32 plane evaluation error relative to the terms.
        assert np.max(np.abs(got - want) / np.maximum(np.abs(want), 1e-3)) < 2.0 ** -18


def test_perspective_correct_interpolation_matches_fp64():
    rng = np.random.default_rng(9)
    for _ in range(40):
        pos = random_triangle(rng)
        col = rng.uniform(-1, 2, (3, 4)).astype(np.float32)
        d = DrawData(width=SIZE, height=SIZE, positions=pos, colors=col, depth_test=0, has_depth=False,
                     color_format=abi.FORMAT_R32G32B32A32_SFLOAT, clear_color=(-9, -9, -9, -9))
        c, _, _, _ = oracle_render(d)
        X, Y, _, W = proj(pos)
        ys, xs = np.nonzero(c[0][..., 0] != -9)
        l0, l1, l2 = barycentrics(X, Y, xs.astype(np.float64), ys.astype(np.float64))
        den = l0 / W[0] + l1 / W[1] + l2 / W[2]
        for k in range(4):
            want = (l0 * col[0, k] / W[0] + l1 * col[1, k] / W[1] + l2 * col[2, k] / W[2]) / den
            got = c[0][ys, xs, k].astype(np.float64)
            assert np.max(np.abs(got - want)) < 2e-4 * (1 + np.abs(col[:, k]).max())


def test_blend_and_unorm_write_within_one_lsb():
    rng = np.random.default_rng(10)
    for _ in range(20):
        pos = random_triangle(rng)
        rgba = rng.uniform(0, 1, 4).astype(np.float32)
        col = np.tile(rgba, (3, 1)).astype(np.float32)
        clear = tuple(rng.uniform(0, 1, 4))
        d = DrawData(width=SIZE, height=SIZE, positions=pos, colors=col, depth_test=0, has_depth=False,
                     clear_color=clear, blend_enable=1,
                     blend=(abi.BLEND_FACTOR_SRC_ALPHA, abi.BLEND_FACTOR_ONE_MINUS_SRC_ALPHA, abi.BLEND_OP_ADD,
                            abi.BLEND_FACTOR_ONE, abi.BLEND_FACTOR_ONE_MINUS_SRC_ALPHA, abi.BLEND_OP_ADD))
        c, _, _, occ = oracle_render(d)
        dst = np.array([int(np.clip(np.rint(np.float32(v) * 255), 0, 255)) for v in clear]) / 255.0
        a = float(rgba[3])
        want = np.empty(4)
        want[:3] = rgba[:3] * a + dst[:3] * (1 - a)
        want[3] = rgba[3] * 1 + dst[3] * (1 - a)
        want = np.clip(want, 0, 1) * 255
        mask = np.any(c[0] != np.round(dst * 255).astype(np.uint8), axis=-1)
        got = c[0][mask].astype(np.float64)
        assert len(got) > 0
        assert np.max(np.abs(got - want)) <= 1.0


def _quad(u0, v0, u1, v1, size=SIZE):
    """Screen-filling quad (two triangles) with w = 1 and the given UV corners."""
    pos = np.array([[-1, -1, 0.5, 1], [1, -1, 0.5, 1], [-1, 1, 0.5, 1],
                    [1, -1, 0.5, 1], [1, 1, 0.5, 1], [-1, 1, 0.5, 1]], np.float32)
    uv = np.array([[u0, v0], [u1, v0], [u0, v1], [u1, v0], [u1, v1], [u0, v1]], np.float32)
    return pos, uv


def _fp64_bilinear(level, u, v):
    h, w = level.shape[:2]
    x = u * w - 0.5
    y = v * h - 0.5
    x0 = np.floor(x).astype(int)
    y0 = np.floor(y).astype(int)
    fx = (x - x0)[..., None]
    fy = (y - y0)[..., None]
    X0, X1 = x0 % w, (x0 + 1) % w
    Y0, Y1 = y0 % h, (y0 + 1) % h
    L = level.astype(np.float64)
    top = L[Y0, X0] * (1 - fx) + L[Y0, X1] * fx
    bot = L[Y1, X0] * (1 - fx) + L[Y1, X1] * fx
    return top * (1 - fy) + bot * fy


@pytest.mark.parametrize("texels_per_pixel,fmt", [(1.0, "f32"), (2.0, "f32"), (2.0 ** 1.5, "f32"),
                                                  (1.0, "u8"), (2.0, "u8"), (2.0 ** 1.5, "u8")])
def test_trilinear_matches_fp64_textbook(texels_per_pixel, fmt):
    """Screen-aligned quad with an exact LOD (L*L a power of two): the sampled
    colour equals textbook wrap-mode trilinear filtering in FP64 within the
    precision of the path (float: 1e-5; 16-bit fixed point: 2/255)."""
    rng = np.random.default_rng(11)
    tex = 64
    if fmt == "f32":
        levels = scenes.float_texture(rng, tex)
        t = TextureData(format=abi.FORMAT_R32G32B32A32_SFLOAT, levels=levels)
        cf = abi.FORMAT_R32G32B32A32_SFLOAT
    else:
        levels = scenes.box_mip_chain(scenes.noise_texture_rgba8(rng, tex))
        t = TextureData(format=abi.FORMAT_R8G8B8A8_UNORM, levels=levels)
        cf = abi.FORMAT_R32G32B32A32_SFLOAT
    span = SIZE * texels_per_pixel / tex  # UV range across the screen
    u0 = 0.1234
    pos, uv = _quad(u0, u0, u0 + span, u0 + span)
    d = DrawData(width=SIZE, height=SIZE, positions=pos, texcoords=uv, shader=abi.SHADER_SAMPLE, texture=t,
                 depth_test=0, has_depth=False, color_format=cf)
    c, _, _, _ = oracle_render(d)
    lod = np.log2(texels_per_pixel)
    ys, xs = np.mgrid[0:SIZE, 0:SIZE]
    u = u0 + (xs + 0.5) / SIZE * span
    v = u0 + (ys + 0.5) / SIZE * span
    l0 = int(np.floor(lod))
    frac = lod - l0
    scale = 1.0 if fmt == "f32" else 255.0
    want = _fp64_bilinear(levels[l0], u, v) / scale
    if frac > 0:
        want = want * (1 - frac) + _fp64_bilinear(levels[l0 + 1], u, v) / scale * frac
    tol = 1e-5 if fmt == "f32" else 2.0 / 255.0
    assert np.max(np.abs(c[0].astype(np.float64) - want)) < tol


def test_constant_texture_is_exact_through_all_filters():
    # Any convex filter of a constant RGBA8 texture returns the constant:
    # exercises weights summing to one in the 16-bit path (incl. aniso).
    v = np.array([200, 17, 255, 0], np.uint8)
    lvl = np.broadcast_to(v, (32, 32, 4)).copy()
    rng = np.random.default_rng(12)
    for aniso in (0, 1):
        t = TextureData(format=abi.FORMAT_R8G8B8A8_UNORM, levels=scenes.box_mip_chain(lvl),
                        anisotropy_enable=aniso, max_anisotropy=8.0)
        pos, uv, col = scenes.screen_triangles(rng, 200, SIZE, SIZE, area_hi=800)
        d = DrawData(width=SIZE, height=SIZE, positions=pos, texcoords=uv, shader=abi.SHADER_SAMPLE, texture=t,
                     color_format=abi.FORMAT_R32G32B32A32_SFLOAT, clear_color=(-1, -1, -1, -1))
        c, _, _, _ = oracle_render(d)
        m = c[0][..., 0] != -1
        got = c[0][m]
        # 16-bit path: c<<8 weighted by 0.16 weights summing to 0xFFFF, scaled by 1/0xFF00.
        assert np.max(np.abs(got - v / 255.0)) <= 1.5 / 255.0
        assert len(got) > 100


def test_singular_uv_jacobian_takes_table_row_zero():
    """With v constant per triangle the anisotropic LOD's determinant is +0,
    the reference's Abs turns it into -0 and anisotropy into -inf; RoundInt
    gives INT_MIN and the wrapped table offsets read row 0 (one probe,
    weight 0xFFFF).  The oracle renders such scenes on both texture paths
    with finite output (before the row was wrapped it read the tables out of
    bounds); the GPU parity scenes of the same name check the bits."""
    from cases import _singular_uv_aniso
    d = _singular_uv_aniso()
    c, _, _, occ = oracle_render(d)
    assert occ > 1000
    f = _singular_uv_aniso(seed=32, float_tex=True, samples=4)
    cf, _, _, occf = oracle_render(f)
    assert occf > 1000 and np.isfinite(cf).all()
