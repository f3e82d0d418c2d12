"""Parity scenes shared by the CPU and GPU tests (seeded, small sizes)."""
import numpy as np

from swiftshader_amd import abi, scenes
from swiftshader_amd.scenes import DrawData, TextureData


def _clip_heavy(seed=11, n=400, size=96):
    """Triangles with vertices well outside the frustum (near/far/sides)."""
    rng = np.random.default_rng(seed)
    xy = rng.uniform(-2.5, 2.5, (n * 3, 2))
    w = rng.uniform(-0.5, 3.0, (n * 3, 1))
    z = rng.uniform(-0.8, 1.6, (n * 3, 1))
    pos = np.concatenate([xy * np.abs(w), z * np.abs(w), w], axis=1).astype(np.float32)
    col = rng.uniform(0, 1, (n * 3, 4)).astype(np.float32)
    uv = rng.uniform(-1, 3, (n * 3, 2)).astype(np.float32)
    tex = TextureData(format=abi.FORMAT_R8G8B8A8_UNORM,
                      levels=scenes.box_mip_chain(scenes.noise_texture_rgba8(rng, 64)))
    return DrawData(width=size, height=size - 17, positions=pos, colors=col, texcoords=uv,
                    shader=abi.SHADER_MODULATE, texture=tex)


def _stencil_scene(seed=12, size=80, samples=1):
    rng = np.random.default_rng(seed)
    pos, uv, col = scenes.screen_triangles(rng, 600, size, size, area_hi=800)
    d = DrawData(width=size, height=size, positions=pos, colors=col, texcoords=uv, samples=samples,
                 shader=abi.SHADER_INTERPOLATE, has_stencil=True, stencil_test=1, clear_stencil=3,
                 depth_compare=abi.COMPARE_LESS_OR_EQUAL,
                 front=dict(fail_op=abi.STENCIL_OP_INCREMENT_AND_WRAP, pass_op=abi.STENCIL_OP_INCREMENT_AND_CLAMP,
                            depth_fail_op=abi.STENCIL_OP_INVERT, compare_op=abi.COMPARE_LESS_OR_EQUAL,
                            compare_mask=0xFE, write_mask=0xF7, reference=2),
                 back=dict(fail_op=abi.STENCIL_OP_DECREMENT_AND_WRAP, pass_op=abi.STENCIL_OP_REPLACE,
                           depth_fail_op=abi.STENCIL_OP_ZERO, compare_op=abi.COMPARE_NOT_EQUAL,
                           compare_mask=0xFF, write_mask=0xFF, reference=5))
    return d


def _blend_variety(seed=13, size=72):
    rng = np.random.default_rng(seed)
    pos, uv, col = scenes.screen_triangles(rng, 500, size, size + 9, area_hi=600)
    col = (col * 1.4 - 0.2).astype(np.float32)  # out-of-range colours exercise the clamps
    return DrawData(width=size, height=size + 9, positions=pos, colors=col, texcoords=uv,
                    shader=abi.SHADER_INTERPOLATE, blend_enable=1,
                    blend=(abi.BLEND_FACTOR_ONE_MINUS_DST_COLOR, abi.BLEND_FACTOR_SRC_ALPHA_SATURATE,
                           abi.BLEND_OP_REVERSE_SUBTRACT, abi.BLEND_FACTOR_CONSTANT_ALPHA,
                           abi.BLEND_FACTOR_ONE_MINUS_CONSTANT_COLOR, abi.BLEND_OP_ADD),
                    blend_constants=(0.25, 1.5, -0.5, 0.75), color_write_mask=0xB,
                    clear_color=(0.2, 0.4, 0.6, 0.8), depth_compare=abi.COMPARE_GREATER_OR_EQUAL,
                    clear_depth=0.0)


def _float_target(seed=14, size=96, tris=1500, samples=1, aniso=False, address=abi.ADDRESS_REPEAT):
    rng = np.random.default_rng(seed)
    tex = TextureData(format=abi.FORMAT_R32G32B32A32_SFLOAT, levels=scenes.float_texture(rng, 128),
                      anisotropy_enable=1 if aniso else 0, max_anisotropy=8.0 if aniso else 1.0,
                      address_u=address, address_v=address)
    pos, uv, col = scenes.screen_triangles(rng, tris, size, size - 8, area_hi=2000)
    return DrawData(width=size, height=size - 8, positions=pos, colors=col, texcoords=uv, samples=samples,
                    shader=abi.SHADER_MODULATE, texture=tex, color_format=abi.FORMAT_R32G32B32A32_SFLOAT,
                    blend_enable=1, blend=(abi.BLEND_FACTOR_SRC_ALPHA, abi.BLEND_FACTOR_ONE_MINUS_SRC_ALPHA,
                                           abi.BLEND_OP_ADD, abi.BLEND_FACTOR_ONE, abi.BLEND_FACTOR_ONE,
                                           abi.BLEND_OP_MAX))


def _rgba8_texture_case(seed=15, size=(160, 90), tris=3000, tex=128, address=abi.ADDRESS_REPEAT, samples=1,
                        aniso=False, mag=abi.FILTER_LINEAR, minf=abi.FILTER_LINEAR, mip=abi.MIPMAP_MODE_LINEAR,
                        npot=None, shader=abi.SHADER_MODULATE, bias=0.0):
    rng = np.random.default_rng(seed)
    w, h = size
    tw = npot[0] if npot else tex
    th = npot[1] if npot else tex
    t = TextureData(format=abi.FORMAT_R8G8B8A8_UNORM, levels=scenes.box_mip_chain(scenes.noise_texture_rgba8(rng, tw, th)),
                    address_u=address, address_v=address, anisotropy_enable=1 if aniso else 0,
                    max_anisotropy=8.0 if aniso else 1.0, mag_filter=mag, min_filter=minf, mipmap_mode=mip,
                    mip_lod_bias=bias)
    pos, uv, col = scenes.screen_triangles(rng, tris, w, h, area_hi=3000)
    return DrawData(width=w, height=h, positions=pos, colors=col, texcoords=uv, samples=samples, shader=shader,
                    texture=t, blend_enable=1,
                    blend=(abi.BLEND_FACTOR_SRC_ALPHA, abi.BLEND_FACTOR_ONE_MINUS_SRC_ALPHA, abi.BLEND_OP_ADD,
                           abi.BLEND_FACTOR_SRC_ALPHA, abi.BLEND_FACTOR_ONE_MINUS_SRC_ALPHA, abi.BLEND_OP_ADD))


def _scissor_cull(seed=16):
    rng = np.random.default_rng(seed)
    pos, uv, col = scenes.screen_triangles(rng, 800, 101, 77, area_hi=1500)
    return DrawData(width=101, height=77, positions=pos, colors=col, shader=abi.SHADER_INTERPOLATE,
                    cull_mode=abi.CULL_BACK, front_face=abi.FRONT_FACE_CLOCKWISE, scissor=(7, 5, 83, 61),
                    viewport=(-3.0, 2.5, 107.0, 70.0, 0.1, 0.9))


def _indexed(seed=17, size=64):
    """A tessellated grid (shared edges) drawn with an index buffer."""
    rng = np.random.default_rng(seed)
    g = 9
    ys, xs = np.meshgrid(np.linspace(-1.1, 1.05, g), np.linspace(-1.02, 1.1, g), indexing="ij")
    jitter = rng.uniform(-0.04, 0.04, (g, g, 2))
    x = xs + jitter[..., 0]
    y = ys + jitter[..., 1]
    z = rng.uniform(0.1, 0.9, (g, g))
    w = rng.uniform(0.7, 1.6, (g, g))
    pos = np.stack([x * w, y * w, z * w, w], -1).reshape(-1, 4).astype(np.float32)
    col = rng.uniform(0, 1, (g * g, 4)).astype(np.float32)
    idx = []
    for i in range(g - 1):
        for j in range(g - 1):
            a, b, c, d = i * g + j, i * g + j + 1, (i + 1) * g + j, (i + 1) * g + j + 1
            idx += [a, b, c, b, d, c]
    return DrawData(width=size, height=size, positions=pos, colors=col, indices=np.asarray(idx, np.uint32),
                    shader=abi.SHADER_INTERPOLATE, depth_compare=abi.COMPARE_ALWAYS)


def _strip_band(seed=18, size=64, rows=6, provoking=abi.PROVOKING_VERTEX_FIRST, indexed=True, clip=False):
    """Horizontal bands drawn as one triangle strip (degenerate joins between
    rows), jittered vertices with perspective w; optionally reaching past
    the frustum so that clipped polygons start at rotated vertices."""
    rng = np.random.default_rng(seed)
    cols = 10
    span = 1.6 if clip else 1.05
    xs = np.linspace(-span, span, cols)
    ys = np.linspace(-span, span, rows + 1)
    grid = np.stack(np.meshgrid(xs, ys, indexing="xy"), -1).reshape(-1, 2)  # (rows + 1) x cols, row-major
    grid = grid + rng.uniform(-0.03, 0.03, grid.shape)
    order = []
    for r in range(rows):
        row = []
        for c in range(cols):
            row += [r * cols + c, (r + 1) * cols + c]
        if order:  # degenerate join to the next row
            order += [order[-1], row[0]]
        order += row
    v = grid
    w = rng.uniform(0.6, 2.0, len(v))
    z = rng.uniform(-0.2 if clip else 0.05, 0.95, len(v))
    pos = np.stack([v[:, 0] * w, v[:, 1] * w, z * w, w], 1).astype(np.float32)
    col = rng.uniform(0, 1, (len(v), 4)).astype(np.float32)
    uv = rng.uniform(-1, 2, (len(v), 2)).astype(np.float32)
    tex = TextureData(format=abi.FORMAT_R8G8B8A8_UNORM, levels=scenes.box_mip_chain(scenes.noise_texture_rgba8(rng, 32)))
    if indexed:
        idx = np.asarray(order, np.uint32)
    else:
        pos, col, uv = pos[order], col[order], uv[order]
        idx = None
    return DrawData(width=size, height=size - 5, positions=pos, colors=col, texcoords=uv, indices=idx,
                    shader=abi.SHADER_MODULATE, texture=tex, topology=abi.TOPOLOGY_TRIANGLE_STRIP,
                    provoking_vertex=provoking, depth_compare=abi.COMPARE_LESS_OR_EQUAL)


def _fan(seed=19, size=72, n=40, provoking=abi.PROVOKING_VERTEX_LAST, samples=1):
    """A star-shaped polygon drawn as a triangle fan around its centre."""
    rng = np.random.default_rng(seed)
    ang = np.sort(rng.uniform(0, 2 * np.pi, n))
    rad = rng.uniform(0.5, 1.3, n)
    hub = np.array([[0.0, 0.0]])
    rim = np.stack([np.cos(ang) * rad, np.sin(ang) * rad], 1)
    rim = np.concatenate([rim, rim[:1]])
    v = np.concatenate([hub, rim])
    w = rng.uniform(0.7, 1.8, len(v))
    z = rng.uniform(0.1, 0.9, len(v))
    pos = np.stack([v[:, 0] * w, v[:, 1] * w, z * w, w], 1).astype(np.float32)
    col = rng.uniform(0, 1, (len(v), 4)).astype(np.float32)
    return DrawData(width=size, height=size, positions=pos, colors=col, shader=abi.SHADER_INTERPOLATE,
                    topology=abi.TOPOLOGY_TRIANGLE_FAN, provoking_vertex=provoking, samples=samples)


def _list_provoking_last(seed=20):
    d = _clip_heavy(seed=seed, n=300)
    d.provoking_vertex = abi.PROVOKING_VERTEX_LAST
    return d


def _singular_uv_aniso(seed=31, float_tex=False, samples=1):
    """Anisotropic filtering where every triangle's v is constant: the quad's
    UV Jacobian is exactly singular (det = +0 -> -0 through Abs), so the
    reference's anisotropy is -inf and its table offsets wrap to row 0."""
    rng = np.random.default_rng(seed)
    pos, uv, col = scenes.screen_triangles(rng, 300, 80, 64, area_lo=30, area_hi=900)
    uv = uv.reshape(-1, 3, 2)
    uv[:, :, 1] = uv[:, :1, 1]
    uv = uv.reshape(-1, 2).astype(np.float32)
    if float_tex:
        tex = TextureData(format=abi.FORMAT_R32G32B32A32_SFLOAT, levels=scenes.float_texture(rng, 64),
                          anisotropy_enable=1, max_anisotropy=8.0)
        fmt = abi.FORMAT_R32G32B32A32_SFLOAT
    else:
        tex = TextureData(format=abi.FORMAT_R8G8B8A8_UNORM,
                          levels=scenes.box_mip_chain(scenes.noise_texture_rgba8(rng, 64)),
                          anisotropy_enable=1, max_anisotropy=8.0)
        fmt = abi.FORMAT_R8G8B8A8_UNORM
    return DrawData(width=80, height=64, positions=pos, colors=col, texcoords=uv, samples=samples,
                    shader=abi.SHADER_MODULATE, texture=tex, color_format=fmt)


def _depth_bias(seed=33, bias=(800.0, 0.0, 1.5), samples=1, size=88, tris=700):
    """Overlapping triangles of varied slope and depth with a depth bias."""
    rng = np.random.default_rng(seed)
    pos, uv, col = scenes.screen_triangles(rng, tris, size, size - 6, area_hi=900)
    return DrawData(width=size, height=size - 6, positions=pos, colors=col, texcoords=uv, samples=samples,
                    shader=abi.SHADER_INTERPOLATE, depth_bias=bias, depth_compare=abi.COMPARE_LESS_OR_EQUAL)


def _lines(seed=36, size=(88, 70), n=300, topology=abi.TOPOLOGY_LINE_LIST, width=1.0,
           mode=abi.LINE_RASTERIZATION_DEFAULT, samples=1, provoking=abi.PROVOKING_VERTEX_FIRST, clip=False,
           textured=False, float_target=False, stencil=False):
    """Random line segments (a list, or one strip through random points); a
    list's first quarter is axis-aligned on the pixel grid (edges through
    pixel centres and between them)."""
    rng = np.random.default_rng(seed)
    W, H = size
    m = 2 * n if topology == abi.TOPOLOGY_LINE_LIST else n + 1
    if clip:
        xy = rng.uniform(-2.0, 2.0, (m, 2))
        w = rng.uniform(-0.4, 2.5, (m, 1))
        z = rng.uniform(-0.6, 1.5, (m, 1))
    else:
        xy = rng.uniform(-1.1, 1.1, (m, 2))
        w = rng.uniform(0.6, 2.0, (m, 1))
        z = rng.uniform(0.05, 0.95, (m, 1))
    if topology == abi.TOPOLOGY_LINE_LIST and not clip:
        k = n // 4
        px = rng.integers(0, 2 * W + 1, (k, 2)) * 0.5  # pixel coordinates on a half-pixel grid
        py = rng.integers(0, 2 * H + 1, (k, 2)) * 0.5
        vert = rng.uniform(size=k) < 0.5
        py[vert, 1] = rng.integers(0, 2 * H + 1, vert.sum()) * 0.5
        px[vert, 1] = px[vert, 0]
        py[~vert, 1] = py[~vert, 0]
        seg = np.stack([px * 2.0 / W - 1.0, py * 2.0 / H - 1.0], -1).reshape(-1, 2)  # (2k, 2) in NDC
        xy[: 2 * k] = seg
        w[: 2 * k] = 1.0
    pos = np.concatenate([xy * np.abs(w), z * np.abs(w), w], axis=1).astype(np.float32)
    col = rng.uniform(0, 1, (m, 4)).astype(np.float32)
    uv = rng.uniform(-1, 3, (m, 2)).astype(np.float32)
    tex = None
    shader = abi.SHADER_INTERPOLATE
    if textured:
        tex = TextureData(format=abi.FORMAT_R8G8B8A8_UNORM,
                          levels=scenes.box_mip_chain(scenes.noise_texture_rgba8(rng, 64)))
        shader = abi.SHADER_MODULATE
    extra = {}
    if stencil:
        extra = dict(has_stencil=True, stencil_test=1, clear_stencil=1,
                     front=dict(fail_op=abi.STENCIL_OP_KEEP, pass_op=abi.STENCIL_OP_INCREMENT_AND_CLAMP,
                                depth_fail_op=abi.STENCIL_OP_INVERT, compare_op=abi.COMPARE_ALWAYS,
                                compare_mask=0xFF, write_mask=0xFF, reference=0),
                     back=dict(fail_op=abi.STENCIL_OP_ZERO, pass_op=abi.STENCIL_OP_ZERO,
                               depth_fail_op=abi.STENCIL_OP_ZERO, compare_op=abi.COMPARE_NEVER,
                               compare_mask=0xFF, write_mask=0xFF, reference=0))
    return DrawData(width=W, height=H, positions=pos, colors=col, texcoords=uv, samples=samples, shader=shader,
                    texture=tex, topology=topology, line_width=width, line_rasterization_mode=mode,
                    provoking_vertex=provoking, depth_compare=abi.COMPARE_LESS_OR_EQUAL,
                    color_format=abi.FORMAT_R32G32B32A32_SFLOAT if float_target else abi.FORMAT_R8G8B8A8_UNORM,
                    **extra)


def _points(seed=41, size=(80, 64), n=700, samples=1, sizes=True, indexed=False, textured=False, clip=False,
            stencil=False):
    """Random points: sizes spread over [0.25, 14] with a few below 1, above
    the target and NaN; optionally indexed, textured, outside the frustum."""
    rng = np.random.default_rng(seed)
    W, H = size
    if clip:
        xy = rng.uniform(-1.6, 1.6, (n, 2))
        w = rng.uniform(-0.3, 2.5, (n, 1))
        z = rng.uniform(-0.5, 1.5, (n, 1))
    else:
        xy = rng.uniform(-1.05, 1.05, (n, 2))
        w = rng.uniform(0.4, 3.0, (n, 1))
        z = rng.uniform(0.05, 0.95, (n, 1))
    pos = np.concatenate([xy * np.abs(w), z * np.abs(w), w], axis=1).astype(np.float32)
    col = rng.uniform(0, 1, (n, 4)).astype(np.float32)
    uv = rng.uniform(-1, 3, (n, 2)).astype(np.float32)
    ps = None
    if sizes:
        ps = rng.uniform(0.25, 14.0, n).astype(np.float32)
        ps[rng.integers(0, n, 12)] = np.float32(np.nan)
        ps[rng.integers(0, n, 3)] = np.float32(2000.0)
        ps[rng.integers(0, n, 20)] = rng.integers(1, 6, 20).astype(np.float32)  # whole sizes: edges on the grid
    idx = None
    if indexed:
        idx = rng.integers(0, n, n + n // 2).astype(np.uint32)
    tex, shader = None, abi.SHADER_INTERPOLATE
    if textured:
        tex = TextureData(format=abi.FORMAT_R8G8B8A8_UNORM,
                          levels=scenes.box_mip_chain(scenes.noise_texture_rgba8(rng, 64)))
        shader = abi.SHADER_MODULATE
    extra = {}
    if stencil:
        extra = dict(has_stencil=True, stencil_test=1, clear_stencil=2,
                     front=dict(fail_op=abi.STENCIL_OP_KEEP, pass_op=abi.STENCIL_OP_INCREMENT_AND_WRAP,
                                depth_fail_op=abi.STENCIL_OP_DECREMENT_AND_CLAMP, compare_op=abi.COMPARE_ALWAYS,
                                compare_mask=0xFF, write_mask=0xFF, reference=0),
                     back=dict(compare_op=abi.COMPARE_NEVER, fail_op=abi.STENCIL_OP_ZERO))
    return DrawData(width=W, height=H, positions=pos, colors=col, texcoords=uv, indices=idx, samples=samples,
                    shader=shader, texture=tex, topology=abi.TOPOLOGY_POINT_LIST, point_sizes=ps,
                    depth_compare=abi.COMPARE_LESS_OR_EQUAL, **extra)


def _formats(seed=44, d16=True, bgra=False, samples=1, bias=None, stencil=False, textured=False, blend=False,
             srgb=False):
    """Overlapping triangles into a D16 depth and/or B8G8R8A8 colour
    attachment (optionally blended with a partial write mask)."""
    rng = np.random.default_rng(seed)
    W, H = 76, 58
    pos, uv, col = scenes.screen_triangles(rng, 900, W, H, area_hi=700)
    col = (col * 1.2 - 0.1).astype(np.float32)
    tex, shader = None, abi.SHADER_INTERPOLATE
    if textured:
        tex = TextureData(format=abi.FORMAT_R8G8B8A8_UNORM,
                          levels=scenes.box_mip_chain(scenes.noise_texture_rgba8(rng, 64)), anisotropy_enable=1,
                          max_anisotropy=4.0)
        shader = abi.SHADER_MODULATE
    extra = {}
    if blend:
        extra.update(blend_enable=1, color_write_mask=0xD,
                     blend=(abi.BLEND_FACTOR_SRC_ALPHA, abi.BLEND_FACTOR_ONE_MINUS_SRC_ALPHA, abi.BLEND_OP_ADD,
                            abi.BLEND_FACTOR_ONE_MINUS_DST_ALPHA, abi.BLEND_FACTOR_ONE, abi.BLEND_OP_ADD),
                     clear_color=(0.7, 0.2, 0.9, 0.4))
    if stencil:
        extra.update(has_stencil=True, stencil_test=1, clear_stencil=9,
                     front=dict(pass_op=abi.STENCIL_OP_INCREMENT_AND_CLAMP, depth_fail_op=abi.STENCIL_OP_DECREMENT_AND_WRAP,
                                compare_op=abi.COMPARE_GREATER, reference=40),
                     back=dict(pass_op=abi.STENCIL_OP_INVERT, depth_fail_op=abi.STENCIL_OP_REPLACE,
                               compare_op=abi.COMPARE_ALWAYS, reference=3))
    return DrawData(width=W, height=H, positions=pos, colors=col, texcoords=uv, samples=samples, shader=shader,
                    texture=tex, depth_format=abi.FORMAT_D16_UNORM if d16 else abi.FORMAT_D32_SFLOAT,
                    color_format=((abi.FORMAT_B8G8R8A8_SRGB if bgra else abi.FORMAT_R8G8B8A8_SRGB) if srgb else
                                  (abi.FORMAT_B8G8R8A8_UNORM if bgra else abi.FORMAT_R8G8B8A8_UNORM)),
                    depth_compare=abi.COMPARE_LESS, depth_bias=bias, clear_depth=0.75, **extra)


def bc_texture_levels(rng, size, n=1, flag=0):
    """A mip chain of random BCn blocks (n = 1..3) and its B8G8R8A8
    decoding by the oracle: [(blocks, decoded (h, w, 4) uint8)] per level."""
    from oracle_runner import bc_block_bytes, oracle_decode_bc
    out = []
    w = h = size
    while True:
        nb = ((w + 3) // 4) * ((h + 3) // 4)
        blocks = rng.integers(0, 256, nb * bc_block_bytes(n), dtype=np.uint8)
        dec = np.zeros((h, w, 4), np.uint8)
        oracle_decode_bc(blocks, dec.reshape(-1), w, h, w * 4, 4, n, flag)
        out.append((blocks, dec))
        if w == 1 and h == 1:
            return out
        w, h = max(w // 2, 1), max(h // 2, 1)


def _bc_textured(seed=51, n=1, aniso=False, samples=1):
    """Triangles sampling a decoded BCn texture (B8G8R8A8 texels)."""
    rng = np.random.default_rng(seed)
    levels = [dec for _, dec in bc_texture_levels(rng, 64, n)]
    tex = TextureData(format=abi.FORMAT_B8G8R8A8_UNORM, levels=levels, anisotropy_enable=1 if aniso else 0,
                      max_anisotropy=8.0 if aniso else 1.0)
    pos, uv, col = scenes.screen_triangles(rng, 1200, 96, 72, area_hi=1500)
    return DrawData(width=96, height=72, positions=pos, colors=col, texcoords=uv, samples=samples,
                    shader=abi.SHADER_MODULATE, texture=tex)


def _advanced_blend(op, k):
    """Overlapping translucent triangles blended with one advanced mode;
    the target format and sample count rotate with k."""
    rng = np.random.default_rng(60 + k)
    W, H = 48, 40
    pos, uv, col = scenes.screen_triangles(rng, 260, W, H, area_hi=500)
    col = col.astype(np.float32)
    col[:, 3] = rng.uniform(0.0, 1.0, len(col)).astype(np.float32)
    col[::7, 3] = 0.0  # zero alpha: colours divide to zero
    if k % 3 == 2:
        col = (col * 1.3 - 0.1).astype(np.float32)  # out of range on a float target
    fmt = (abi.FORMAT_R8G8B8A8_UNORM, abi.FORMAT_B8G8R8A8_SRGB, abi.FORMAT_R32G32B32A32_SFLOAT)[k % 3]
    return DrawData(width=W, height=H, positions=pos, colors=col, shader=abi.SHADER_INTERPOLATE,
                    samples=4 if k % 4 == 1 else 1, color_format=fmt, blend_enable=1,
                    blend=(abi.BLEND_FACTOR_SRC_ALPHA, abi.BLEND_FACTOR_ZERO, op, abi.BLEND_FACTOR_ONE,
                           abi.BLEND_FACTOR_ZERO, op),
                    clear_color=(0.2, 0.5, 0.8, 0.6), depth_compare=abi.COMPARE_LESS_OR_EQUAL)


def _srgb_textured(seed=53, bgra=False, aniso=False, samples=1, target=abi.FORMAT_R8G8B8A8_UNORM):
    """Triangles sampling an sRGB-encoded texture (decoded per texel before
    filtering)."""
    rng = np.random.default_rng(seed)
    levels = scenes.box_mip_chain(scenes.noise_texture_rgba8(rng, 64))
    if bgra:
        levels = [np.ascontiguousarray(l[..., [2, 1, 0, 3]]) for l in levels]
    tex = TextureData(format=abi.FORMAT_B8G8R8A8_SRGB if bgra else abi.FORMAT_R8G8B8A8_SRGB, levels=levels,
                      anisotropy_enable=1 if aniso else 0, max_anisotropy=8.0 if aniso else 1.0)
    pos, uv, col = scenes.screen_triangles(rng, 1000, 88, 66, area_hi=1500)
    return DrawData(width=88, height=66, positions=pos, colors=col, texcoords=uv, samples=samples,
                    shader=abi.SHADER_MODULATE, texture=tex, color_format=target)


def _with_index_type(draw, dtype):
    """The same indexed draw with its indices narrowed to uint16 or uint8
    (VkIndexType UINT16 / UINT8_EXT)."""
    assert draw.indices.max() <= np.iinfo(dtype).max
    draw.indices = draw.indices.astype(dtype)
    return draw


def _depth_clamped(draw, clamp, clip, near=0.15, far=0.8, compare=None):
    """The draw with depth clamp / depth clip state and a viewport depth
    range inside [0, 1], so that clamping to it is visible."""
    draw.depth_clamp_enable, draw.depth_clip_control = clamp, clip
    if compare is not None:
        draw.depth_compare = compare
    draw.viewport = (0.0, 0.0, float(draw.width), float(draw.height), near, far)
    return draw


def _depth_bounded(draw, lo, hi, clear_depth, depth_test=1, depth_format=None):
    """The draw with the depth bounds test on [lo, hi], cleared to a depth
    inside them so that the first triangles can pass."""
    draw.depth_bounds, draw.clear_depth, draw.depth_test = (lo, hi), clear_depth, depth_test
    if depth_format is not None:
        draw.depth_format = depth_format
    return draw


def _alpha_to_coverage(seed=70, samples=4, shader=abi.SHADER_MODULATE, float_target=False, stencil=False,
                       blend=False, size=(72, 56), tris=700):
    """Alpha-to-coverage over random alphas (some outside [0, 1]), with
    depth LESS, optionally stencil ops, blending and a float target."""
    rng = np.random.default_rng(seed)
    pos, uv, col = scenes.screen_triangles(rng, tris, size[0], size[1], area_hi=900)
    col[:, 3] = rng.uniform(-0.2, 1.2, len(col)).astype(np.float32)
    tex = TextureData(format=abi.FORMAT_R8G8B8A8_UNORM,
                      levels=scenes.box_mip_chain(scenes.noise_texture_rgba8(rng, 64)))
    d = DrawData(width=size[0], height=size[1], positions=pos, colors=col, texcoords=uv, samples=samples,
                 shader=shader, texture=tex if shader != abi.SHADER_INTERPOLATE else None, alpha_to_coverage=1,
                 color_format=abi.FORMAT_R32G32B32A32_SFLOAT if float_target else abi.FORMAT_R8G8B8A8_UNORM)
    if stencil:
        d.has_stencil, d.stencil_test, d.clear_stencil = True, 1, 1
        d.front = dict(fail_op=abi.STENCIL_OP_INVERT, pass_op=abi.STENCIL_OP_INCREMENT_AND_WRAP,
                       depth_fail_op=abi.STENCIL_OP_DECREMENT_AND_CLAMP, compare_op=abi.COMPARE_NOT_EQUAL,
                       compare_mask=0xFF, write_mask=0xFF, reference=6)
        d.back = dict(fail_op=abi.STENCIL_OP_ZERO, pass_op=abi.STENCIL_OP_REPLACE,
                      depth_fail_op=abi.STENCIL_OP_INCREMENT_AND_CLAMP, compare_op=abi.COMPARE_ALWAYS,
                      compare_mask=0xFF, write_mask=0x7F, reference=9)
    if blend:
        d.blend_enable = 1
        d.blend = (abi.BLEND_FACTOR_SRC_ALPHA, abi.BLEND_FACTOR_ONE_MINUS_SRC_ALPHA, abi.BLEND_OP_ADD,
                   abi.BLEND_FACTOR_ONE, abi.BLEND_FACTOR_ONE_MINUS_SRC_ALPHA, abi.BLEND_OP_ADD)
    return d


def _bounds_ramp(samples=1, depth_test=1, depth_format=abi.FORMAT_D32_SFLOAT, bounds=(0.3, 0.7)):
    """A depth ramp (z = (x + 1) / 2) then a flat quad at 1/2 drawn with
    ALWAYS under the depth bounds test, with stencil ops; at 4 samples a
    column's samples straddle the bounds differently."""
    xy = np.array([[-1, -1], [1, -1], [1, 1], [-1, -1], [1, 1], [-1, 1]], np.float64)
    ramp = np.concatenate([xy, (xy[:, :1] + 1) / 2, np.ones((6, 1))], axis=1)
    flat = np.concatenate([xy, np.full((6, 1), 0.5), np.ones((6, 1))], axis=1)
    pos = np.concatenate([ramp, flat]).astype(np.float32)
    col = np.array([[1, 0, 0, 1]] * 6 + [[0, 1, 0, 1]] * 6, np.float32)
    st = dict(fail_op=abi.STENCIL_OP_ZERO, pass_op=abi.STENCIL_OP_REPLACE,
              depth_fail_op=abi.STENCIL_OP_INCREMENT_AND_CLAMP, compare_op=abi.COMPARE_ALWAYS,
              compare_mask=0xFF, write_mask=0xFF, reference=7)
    return DrawData(width=45, height=6, positions=pos, colors=col, samples=samples, depth_test=depth_test,
                    depth_compare=abi.COMPARE_ALWAYS, clear_depth=0.5, depth_format=depth_format,
                    has_stencil=True, stencil_test=1, front=st, back=st, depth_bounds=bounds)


def parity_cases():
    """(name, DrawData) pairs; every one must match the oracle bit for bit."""
    return [
        ("plain_gouraud_256", scenes.plain_gouraud()),
        ("textured_small", scenes.textured_1080p(triangles=3000, width=192, height=108, tex_size=128)),
        ("rgba8_clamp_npot", _rgba8_texture_case(address=abi.ADDRESS_CLAMP_TO_EDGE, npot=(100, 37), seed=21)),
        ("rgba8_nearest_mips", _rgba8_texture_case(mag=abi.FILTER_NEAREST, minf=abi.FILTER_NEAREST,
                                                  mip=abi.MIPMAP_MODE_NEAREST, seed=22, shader=abi.SHADER_SAMPLE)),
        ("rgba8_mixed_filters", _rgba8_texture_case(mag=abi.FILTER_NEAREST, minf=abi.FILTER_LINEAR, seed=23,
                                                   bias=0.7)),
        ("rgba8_aniso", _rgba8_texture_case(aniso=True, seed=24)),
        ("rgba8_msaa4_aniso", _rgba8_texture_case(aniso=True, samples=4, seed=25, size=(70, 50), tris=900)),
        ("float_target", _float_target()),
        ("float_msaa4_aniso", _float_target(samples=4, aniso=True, seed=26, size=64, tris=700)),
        ("float_clamp_aniso", _float_target(aniso=True, seed=27, address=abi.ADDRESS_CLAMP_TO_EDGE)),
        ("clip_heavy", _clip_heavy()),
        ("stencil_ops", _stencil_scene()),
        ("stencil_msaa4", _stencil_scene(seed=28, samples=4)),
        ("blend_variety", _blend_variety()),
        ("scissor_cull_viewport", _scissor_cull()),
        ("indexed_grid", _indexed()),
        ("strip_indexed", _strip_band()),
        ("strip_provoking_last_clipped", _strip_band(seed=29, provoking=abi.PROVOKING_VERTEX_LAST, indexed=False,
                                                     clip=True)),
        ("fan_provoking_last", _fan()),
        ("fan_msaa4", _fan(seed=30, provoking=abi.PROVOKING_VERTEX_FIRST, samples=4)),
        ("list_provoking_last_clipped", _list_provoking_last()),
        ("singular_uv_aniso_rgba8", _singular_uv_aniso()),
        ("singular_uv_aniso_float_msaa4", _singular_uv_aniso(seed=32, float_tex=True, samples=4)),
        ("depth_bias_constant_slope", _depth_bias()),
        ("depth_bias_clamp_negative_msaa4", _depth_bias(seed=34, bias=(-3000.0, -0.002, -2.0), samples=4)),
        ("depth_bias_clamp_positive", _depth_bias(seed=35, bias=(1e6, 0.004, 0.0))),
        ("lines_list_gouraud", _lines()),
        ("lines_strip_wide_msaa4_textured", _lines(seed=37, topology=abi.TOPOLOGY_LINE_STRIP, n=120, width=3.5,
                                                   samples=4, textured=True, provoking=abi.PROVOKING_VERTEX_LAST)),
        ("lines_bresenham_float", _lines(seed=38, mode=abi.LINE_RASTERIZATION_BRESENHAM, float_target=True)),
        ("lines_bresenham_wide_strip", _lines(seed=39, topology=abi.TOPOLOGY_LINE_STRIP, n=150, width=2.5,
                                              mode=abi.LINE_RASTERIZATION_BRESENHAM, textured=True)),
        ("lines_clipped_last_msaa4_stencil", _lines(seed=40, clip=True, width=1.5, samples=4, stencil=True,
                                                    provoking=abi.PROVOKING_VERTEX_LAST)),
        ("points_sizes", _points()),
        ("points_sizes_msaa4_clipped_stencil", _points(seed=42, samples=4, clip=True, stencil=True)),
        ("points_indexed_textured_unsized", _points(seed=43, sizes=False, indexed=True, textured=True)),
        ("d16_depth", _formats()),
        ("d16_msaa4_bias_stencil", _formats(seed=45, samples=4, bias=(-3.0, 0.0, 1.25), stencil=True)),
        ("bgra8_blend_mask", _formats(seed=46, d16=False, bgra=True, blend=True)),
        ("bgra8_d16_msaa4_textured", _formats(seed=47, bgra=True, samples=4, textured=True)),
        ("srgb8_blend_mask", _formats(seed=48, d16=False, srgb=True, blend=True)),
        ("bgra8_srgb_msaa4_textured_blend", _formats(seed=49, bgra=True, srgb=True, samples=4, textured=True,
                                                    blend=True)),
        ("srgb8_opaque_d16", _formats(seed=50, srgb=True)),
        ("bc1_texture_bgra", _bc_textured()),
        ("bc3_texture_bgra_aniso_msaa4", _bc_textured(seed=52, n=3, aniso=True, samples=4)),
        ("indexed_grid_uint16", _with_index_type(_indexed(seed=55), np.uint16)),
        ("strip_indexed_uint8", _with_index_type(_strip_band(seed=56), np.uint8)),
        ("depth_clamp_clip_heavy", _depth_clamped(_clip_heavy(seed=59), clamp=1, clip=0,
                                                  compare=abi.COMPARE_ALWAYS)),
        ("depth_clip_off_lines_msaa4", _depth_clamped(_lines(seed=61, clip=True, samples=4), clamp=0, clip=2)),
        ("depth_clamp_clip_on_points", _depth_clamped(_points(seed=62, clip=True), clamp=1, clip=1)),
        ("depth_bounds_gouraud", _depth_bounded(scenes.plain_gouraud(), 0.2, 0.6, clear_depth=0.55)),
        ("depth_bounds_d16_textured",
         _depth_bounded(scenes.textured_1080p(triangles=3000, width=192, height=108, tex_size=128), 0.1, 0.5,
                        clear_depth=0.45, depth_format=abi.FORMAT_D16_UNORM)),
        ("depth_bounds_depth_test_off", _depth_bounded(scenes.plain_gouraud(), 0.6, 0.9, clear_depth=0.5,
                                                       depth_test=0)),
        ("rgba8_mirrored_npot_aniso", _rgba8_texture_case(address=abi.ADDRESS_MIRRORED_REPEAT, npot=(100, 37),
                                                          aniso=True, seed=63)),
        ("rgba8_mirror_clamp_msaa4", _rgba8_texture_case(address=abi.ADDRESS_MIRROR_CLAMP_TO_EDGE, samples=4, seed=64,
                                                         size=(70, 50), tris=900)),
        ("float_mirrored_aniso", _float_target(address=abi.ADDRESS_MIRRORED_REPEAT, aniso=True, seed=65)),
        ("float_mirror_clamp", _float_target(address=abi.ADDRESS_MIRROR_CLAMP_TO_EDGE, seed=66)),
        ("srgb_texture_trilinear", _srgb_textured()),
        ("bgra_srgb_texture_aniso_msaa4_srgb_target", _srgb_textured(seed=54, bgra=True, aniso=True, samples=4,
                                                                     target=abi.FORMAT_R8G8B8A8_SRGB)),
        ("a2c_msaa4_modulate", _alpha_to_coverage()),
        ("a2c_msaa4_stencil_blend", _alpha_to_coverage(seed=71, stencil=True, blend=True)),
        ("a2c_1x_gouraud", _alpha_to_coverage(seed=72, samples=1, shader=abi.SHADER_INTERPOLATE)),
        ("a2c_msaa4_float_target", _alpha_to_coverage(seed=73, float_target=True, shader=abi.SHADER_INTERPOLATE)),
        ("depth_bounds_ramp_stencil_d32f", _bounds_ramp()),
        ("depth_bounds_ramp_stencil_d16", _bounds_ramp(depth_format=abi.FORMAT_D16_UNORM)),
        ("depth_bounds_ramp_stencil_test_off", _bounds_ramp(depth_test=0)),
        ("depth_bounds_ramp_stencil_msaa4", _bounds_ramp(samples=4)),
        ("depth_bounds_ramp_stencil_msaa4_d16_test_off", _bounds_ramp(samples=4, depth_test=0,
                                                                     depth_format=abi.FORMAT_D16_UNORM)),
    ] + [(f"advanced_blend_{op - 1000148000}", _advanced_blend(op, k)) for k, op in enumerate(abi.ADVANCED_BLEND_OPS)]
