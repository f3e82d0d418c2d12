"""Seeded synthetic workloads for the draw path.

Each generator returns plain numpy arrays (identical bytes for the CPU
oracle and the GPU path) describing post-vertex-stage clip-space vertices,
the fixed-function state and the textures of one or more draws.

Workloads (sizes are the ones the benchmark and parity tests use):
  plain_gouraud     256x256 RGBA8 + D32F, 1,000 triangles, x,y ~ U[-1,1],
                    z ~ U[0,1], w = 1, per-vertex RGBA ~ U[0,1], LESS, no blend.
  textured_1080p    1920x1080 RGBA8 + D32F, 100,000 triangles of screen area
                    log-uniform in [1, 1e4] px, w ~ U[0.5,4], UV ~ U[0,4],
                    1024^2 RGBA8 noise texture with an 11-level box mip chain,
                    trilinear, REPEAT, LESS, SRC_ALPHA/ONE_MINUS_SRC_ALPHA blend,
                    alpha ~ U[0.2,1], texture x colour.
  batched           512 draws, targets log-uniform 64..1024 per side, 2k..50k
                    triangles, per-draw RGBA8 noise texture of side 256..2048
                    (NPOT), full mips, trilinear, CLAMP_TO_EDGE or REPEAT.
  diff_views        16 views 512^2 RGBA32F + D32F, 20k triangles each, one
                    1024^2 RGBA32F texture ~ N(0.5, 0.15), trilinear.
  msaa_aniso        64 views of side log-uniform 256..1024, 4x MSAA, 10k..40k
                    triangles, 2048^2 RGBA32F texture, 8x anisotropic.
Every generator accepts a `scale` factor that shrinks triangle counts (and
target sizes where stated) for fast parity tests.
"""
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from . import abi


@dataclass
class TextureData:
    format: int
    levels: List[np.ndarray]           # each (h, w, 4) uint8 or float32, C-contiguous
    mag_filter: int = abi.FILTER_LINEAR
    min_filter: int = abi.FILTER_LINEAR
    mipmap_mode: int = abi.MIPMAP_MODE_LINEAR
    address_u: int = abi.ADDRESS_REPEAT
    address_v: int = abi.ADDRESS_REPEAT
    mip_lod_bias: float = 0.0
    anisotropy_enable: int = 0
    max_anisotropy: float = 1.0
    min_lod: float = 0.0
    max_lod: float = 1000.0            # VK_LOD_CLAMP_NONE; clamped to 13
    size: int = 0                      # level-0 side when `levels` is left for the device to fill


@dataclass
class DrawData:
    width: int
    height: int
    positions: np.ndarray               # (V, 4) float32 clip space
    colors: Optional[np.ndarray] = None  # (V, 4) float32
    texcoords: Optional[np.ndarray] = None  # (V, 2) float32
    indices: Optional[np.ndarray] = None    # (T*3,) uint32
    samples: int = 1
    sample_mask: int = 0xFFFFFFFF
    color_format: int = abi.FORMAT_R8G8B8A8_UNORM
    depth_format: int = abi.FORMAT_D32_SFLOAT  # or FORMAT_D16_UNORM
    shader: int = abi.SHADER_INTERPOLATE
    cull_mode: int = abi.CULL_NONE
    front_face: int = abi.FRONT_FACE_COUNTER_CLOCKWISE
    depth_test: int = 1
    depth_write: int = 1
    depth_compare: int = abi.COMPARE_LESS
    stencil_test: int = 0
    front: dict = field(default_factory=dict)
    back: dict = field(default_factory=dict)
    blend_enable: int = 0
    blend: tuple = (abi.BLEND_FACTOR_ONE, abi.BLEND_FACTOR_ZERO, abi.BLEND_OP_ADD,
                    abi.BLEND_FACTOR_ONE, abi.BLEND_FACTOR_ZERO, abi.BLEND_OP_ADD)
    color_write_mask: int = 0xF
    blend_constants: tuple = (0.0, 0.0, 0.0, 0.0)
    texture: Optional[TextureData] = None
    scissor: Optional[tuple] = None     # (x, y, w, h); default = full target
    viewport: Optional[tuple] = None    # (x, y, w, h, minD, maxD)
    clear_color: tuple = (0.0, 0.0, 0.0, 0.0)
    clear_depth: float = 1.0
    clear_stencil: int = 0
    has_depth: bool = True
    has_stencil: bool = False
    topology: int = abi.TOPOLOGY_TRIANGLE_LIST
    provoking_vertex: int = abi.PROVOKING_VERTEX_FIRST
    depth_bias: Optional[tuple] = None  # (constant_factor, clamp, slope_factor); None = disabled
    line_width: float = 1.0
    line_rasterization_mode: int = abi.LINE_RASTERIZATION_DEFAULT
    point_sizes: Optional[np.ndarray] = None  # (V,) float32 PointSize per vertex; None = 1
    depth_clamp_enable: int = 0
    depth_clip_control: int = 0  # 0 default (clip unless clamping), 1 clip, 2 no clip
    depth_bounds: Optional[Tuple[float, float]] = None  # (min, max) enables the depth bounds test
    alpha_to_coverage: int = 0

    @property
    def triangle_count(self):
        """Primitives in the draw (lines for the line topologies)."""
        n = len(self.indices) if self.indices is not None else len(self.positions)
        if self.topology in (abi.TOPOLOGY_TRIANGLE_STRIP, abi.TOPOLOGY_TRIANGLE_FAN):
            return max(n - 2, 0)
        if self.topology == abi.TOPOLOGY_LINE_STRIP:
            return max(n - 1, 0)
        if self.topology == abi.TOPOLOGY_LINE_LIST:
            return n // 2
        if self.topology == abi.TOPOLOGY_POINT_LIST:
            return n
        return n // 3


def box_mip_chain(level0: np.ndarray) -> List[np.ndarray]:
    """Full mip chain by 2x2 box filtering (odd sizes: floor, edge clamp)."""
    levels = [np.ascontiguousarray(level0)]
    cur = level0
    is_u8 = level0.dtype == np.uint8
    while cur.shape[0] > 1 or cur.shape[1] > 1:
        h, w = cur.shape[:2]
        nh, nw = max(h // 2, 1), max(w // 2, 1)
        ys0 = np.minimum(2 * np.arange(nh), h - 1)
        ys1 = np.minimum(2 * np.arange(nh) + 1, h - 1)
        xs0 = np.minimum(2 * np.arange(nw), w - 1)
        xs1 = np.minimum(2 * np.arange(nw) + 1, w - 1)
        if is_u8:
            acc = (cur[ys0][:, xs0].astype(np.uint32) + cur[ys0][:, xs1] + cur[ys1][:, xs0] + cur[ys1][:, xs1] + 2) // 4
            nxt = acc.astype(np.uint8)
        else:
            nxt = ((cur[ys0][:, xs0] + cur[ys0][:, xs1]) + (cur[ys1][:, xs0] + cur[ys1][:, xs1])) * np.float32(0.25)
            nxt = nxt.astype(np.float32)
        levels.append(np.ascontiguousarray(nxt))
        cur = nxt
    return levels


def noise_texture_rgba8(rng, size_w, size_h=None):
    size_h = size_w if size_h is None else size_h
    return rng.integers(0, 256, size=(size_h, size_w, 4), dtype=np.uint8)


def screen_triangles(rng, n, width, height, area_lo=1.0, area_hi=1e4, w_lo=0.5, w_hi=4.0,
                     uv_hi=4.0, alpha_lo=0.2):
    """Triangles of log-uniform screen area, lifted to clip space.

    Each triangle is placed in screen space (centre uniform over the target,
    random shape, area log-uniform in [area_lo, area_hi] pixels), mapped to
    NDC, given z ~ U[0,1] and a per-vertex w ~ U[w_lo, w_hi], and multiplied
    by w so that the projection returns the placed vertex.
    """
    area_hi = min(area_hi, 0.5 * width * height)
    area = np.exp(rng.uniform(np.log(area_lo), np.log(area_hi), n))
    cx = rng.uniform(0, width, n)
    cy = rng.uniform(0, height, n)
    ang = rng.uniform(0, 2 * np.pi, (n, 3))
    rad = rng.uniform(0.4, 1.0, (n, 3))
    px = np.cos(ang) * rad
    py = np.sin(ang) * rad
    a = 0.5 * np.abs((px[:, 1] - px[:, 0]) * (py[:, 2] - py[:, 0]) - (px[:, 2] - px[:, 0]) * (py[:, 1] - py[:, 0]))
    a = np.maximum(a, 1e-3)
    s = np.sqrt(area / a)
    sx = cx[:, None] + px * s[:, None]
    sy = cy[:, None] + py * s[:, None]
    ndc_x = 2.0 * sx / width - 1.0
    ndc_y = 2.0 * sy / height - 1.0
    z = rng.uniform(0.0, 1.0, (n, 3))
    w = rng.uniform(w_lo, w_hi, (n, 3))
    pos = np.stack([ndc_x * w, ndc_y * w, z * w, w], axis=-1).reshape(-1, 4).astype(np.float32)
    uv = rng.uniform(0.0, uv_hi, (n * 3, 2)).astype(np.float32)
    col = rng.uniform(0.0, 1.0, (n * 3, 4)).astype(np.float32)
    col[:, 3] = rng.uniform(alpha_lo, 1.0, n * 3).astype(np.float32)
    return pos, uv, col


def plain_gouraud(seed=1, triangles=1000, size=256):
    rng = np.random.default_rng(seed)
    n = triangles
    xy = rng.uniform(-1.0, 1.0, (n * 3, 2))
    z = rng.uniform(0.0, 1.0, (n * 3, 1))
    pos = np.concatenate([xy, z, np.ones((n * 3, 1))], axis=1).astype(np.float32)
    col = rng.uniform(0.0, 1.0, (n * 3, 4)).astype(np.float32)
    return DrawData(width=size, height=size, positions=pos, colors=col, shader=abi.SHADER_INTERPOLATE)


def textured_1080p(seed=2, triangles=100_000, width=1920, height=1080, tex_size=1024):
    rng = np.random.default_rng(seed)
    pos, uv, col = screen_triangles(rng, triangles, width, height)
    tex = TextureData(format=abi.FORMAT_R8G8B8A8_UNORM, levels=box_mip_chain(noise_texture_rgba8(rng, tex_size)))
    return DrawData(width=width, height=height, positions=pos, colors=col, texcoords=uv,
                    shader=abi.SHADER_MODULATE, texture=tex, blend_enable=1,
                    blend=(abi.BLEND_FACTOR_SRC_ALPHA, abi.BLEND_FACTOR_ONE_MINUS_SRC_ALPHA, abi.BLEND_OP_ADD,
                           abi.BLEND_FACTOR_SRC_ALPHA, abi.BLEND_FACTOR_ONE_MINUS_SRC_ALPHA, abi.BLEND_OP_ADD))


def _log_uniform_int(rng, lo, hi):
    return int(round(float(np.exp(rng.uniform(np.log(lo), np.log(hi))))))


def batched(seed=3, draws=512, tri_lo=2000, tri_hi=50000, rt_lo=64, rt_hi=1024, tex_lo=256, tex_hi=2048,
            scale=1.0, host_textures=True):
    """The batched workload.  With host_textures=False the textures' texels
    are not generated here (levels stays empty, TextureData.size holds the
    side): the caller fills them on the device."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(draws):
        w = _log_uniform_int(rng, rt_lo, rt_hi)
        h = _log_uniform_int(rng, rt_lo, rt_hi)
        n = max(1, int(_log_uniform_int(rng, tri_lo, tri_hi) * scale))
        ts = _log_uniform_int(rng, tex_lo, tex_hi)
        pos, uv, col = screen_triangles(rng, n, w, h, area_hi=min(1e4, w * h / 4))
        mode = abi.ADDRESS_CLAMP_TO_EDGE if rng.integers(0, 2) else abi.ADDRESS_REPEAT
        levels = box_mip_chain(noise_texture_rgba8(rng, ts)) if host_textures else []
        tex = TextureData(format=abi.FORMAT_R8G8B8A8_UNORM, levels=levels, address_u=mode, address_v=mode, size=ts)
        out.append(DrawData(width=w, height=h, positions=pos, colors=col, texcoords=uv,
                            shader=abi.SHADER_MODULATE, texture=tex, blend_enable=1,
                            blend=(abi.BLEND_FACTOR_SRC_ALPHA, abi.BLEND_FACTOR_ONE_MINUS_SRC_ALPHA,
                                   abi.BLEND_OP_ADD, abi.BLEND_FACTOR_SRC_ALPHA,
                                   abi.BLEND_FACTOR_ONE_MINUS_SRC_ALPHA, abi.BLEND_OP_ADD)))
    return out


def float_texture(rng, size, mean=0.5, std=0.15):
    lvl0 = rng.normal(mean, std, (size, size, 4)).astype(np.float32)
    return box_mip_chain(lvl0)


def diff_views(seed=4, views=16, triangles=20000, size=512, tex_size=1024, texture=None):
    rng = np.random.default_rng(seed)
    if texture is None:
        texture = TextureData(format=abi.FORMAT_R32G32B32A32_SFLOAT, levels=float_texture(rng, tex_size))
    out = []
    for _ in range(views):
        pos, uv, col = screen_triangles(rng, triangles, size, size, area_hi=min(1e4, size * size / 8))
        out.append(DrawData(width=size, height=size, positions=pos, colors=col, texcoords=uv,
                            shader=abi.SHADER_MODULATE, texture=texture,
                            color_format=abi.FORMAT_R32G32B32A32_SFLOAT))
    return out


def msaa_aniso(seed=5, views=64, tri_lo=10000, tri_hi=40000, side_lo=256, side_hi=1024, tex_size=2048,
               scale=1.0, texture=None):
    rng = np.random.default_rng(seed)
    if texture is None:
        texture = TextureData(format=abi.FORMAT_R32G32B32A32_SFLOAT, levels=float_texture(rng, tex_size),
                              anisotropy_enable=1, max_anisotropy=8.0)
    out = []
    for _ in range(views):
        w = _log_uniform_int(rng, side_lo, side_hi)
        h = _log_uniform_int(rng, side_lo, side_hi)
        n = max(1, int(_log_uniform_int(rng, tri_lo, tri_hi) * scale))
        pos, uv, col = screen_triangles(rng, n, w, h, area_hi=min(1e4, w * h / 8))
        out.append(DrawData(width=w, height=h, positions=pos, colors=col, texcoords=uv, samples=4,
                            shader=abi.SHADER_MODULATE, texture=texture,
                            color_format=abi.FORMAT_R32G32B32A32_SFLOAT))
    return out
