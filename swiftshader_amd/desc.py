"""Builds swr_draw_desc structures from a DrawData and buffer addresses.

The same builder serves the HIP path (device addresses of torch tensors)
and any host caller; it only fills the C struct, it never touches memory.
Field meanings follow the DrawCall / DrawData inputs of the reference's
sw::Renderer::draw (src/Device/Renderer.cpp:183-490).
"""
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from . import abi
from .scenes import DrawData


@dataclass
class Buffers:
    positions: int
    colors: int = 0
    texcoords: int = 0
    indices: int = 0
    point_sizes: int = 0
    texture_levels: Optional[List[Tuple[int, int, int, int]]] = None  # (addr, width, height, row_pitch_bytes)
    color: int = 0
    color_pitch: int = 0
    color_sample_pitch: int = 0
    depth: int = 0
    depth_pitch: int = 0
    depth_sample_pitch: int = 0
    stencil: int = 0
    stencil_pitch: int = 0
    stencil_sample_pitch: int = 0
    occlusion: int = 0


def _stencil_state(d: dict) -> abi.StencilOpState:
    s = abi.StencilOpState()
    s.fail_op = d.get("fail_op", abi.STENCIL_OP_KEEP)
    s.pass_op = d.get("pass_op", abi.STENCIL_OP_KEEP)
    s.depth_fail_op = d.get("depth_fail_op", abi.STENCIL_OP_KEEP)
    s.compare_op = d.get("compare_op", abi.COMPARE_ALWAYS)
    s.compare_mask = d.get("compare_mask", 0xFF)
    s.write_mask = d.get("write_mask", 0xFF)
    s.reference = d.get("reference", 0)
    return s


def index_type_of(indices) -> int:
    """VkIndexType of an index array by its dtype (uint8, uint16, else uint32)."""
    if indices is not None and indices.dtype == np.uint8:
        return abi.INDEX_TYPE_UINT8
    if indices is not None and indices.dtype == np.uint16:
        return abi.INDEX_TYPE_UINT16
    return abi.INDEX_TYPE_UINT32


def color_bytes_per_pixel(fmt: int) -> int:
    return 16 if fmt == abi.FORMAT_R32G32B32A32_SFLOAT else 4


def depth_bytes_per_pixel(fmt: int) -> int:
    return 2 if fmt == abi.FORMAT_D16_UNORM else 4


# Bytes per pixel of every attachment format the library clears.
FORMAT_BYTES = {abi.FORMAT_R8G8B8A8_UNORM: 4, abi.FORMAT_B8G8R8A8_UNORM: 4, abi.FORMAT_R8G8B8A8_SRGB: 4,
                abi.FORMAT_B8G8R8A8_SRGB: 4, abi.FORMAT_R32G32B32A32_SFLOAT: 16, abi.FORMAT_D16_UNORM: 2,
                abi.FORMAT_D32_SFLOAT: 4, abi.FORMAT_S8_UINT: 1}
COLOR_FORMATS = (abi.FORMAT_R8G8B8A8_UNORM, abi.FORMAT_B8G8R8A8_UNORM, abi.FORMAT_R8G8B8A8_SRGB,
                 abi.FORMAT_B8G8R8A8_SRGB, abi.FORMAT_R32G32B32A32_SFLOAT)
DEPTH_FORMATS = (abi.FORMAT_D16_UNORM, abi.FORMAT_D32_SFLOAT)


def build_desc(draw: DrawData, buf: Buffers) -> abi.DrawDesc:
    d = abi.DrawDesc()
    d.positions = buf.positions
    d.colors = buf.colors or None
    d.texcoords = buf.texcoords or None
    d.indices = buf.indices or None
    d.point_sizes = buf.point_sizes or None
    d.index_type = index_type_of(draw.indices)
    d.depth_clamp_enable = draw.depth_clamp_enable
    d.depth_clip_control = draw.depth_clip_control
    if draw.depth_bounds is not None:
        d.depth_bounds_test_enable = 1
        d.min_depth_bounds, d.max_depth_bounds = draw.depth_bounds
    d.alpha_to_coverage_enable = 1 if draw.alpha_to_coverage else 0
    d.vertex_count = len(draw.positions)
    d.triangle_count = draw.triangle_count
    if draw.viewport is None:
        d.viewport = abi.Viewport(0.0, 0.0, float(draw.width), float(draw.height), 0.0, 1.0)
    else:
        d.viewport = abi.Viewport(*[float(v) for v in draw.viewport])
    if draw.scissor is None:
        d.scissor = abi.Rect(0, 0, draw.width, draw.height)
    else:
        d.scissor = abi.Rect(*draw.scissor)
    d.render_area = abi.Rect(0, 0, draw.width, draw.height)
    d.cull_mode = draw.cull_mode
    d.front_face = draw.front_face
    d.samples = draw.samples
    d.sample_mask = draw.sample_mask & 0xFFFFFFFF
    d.depth_test_enable = draw.depth_test
    d.depth_write_enable = draw.depth_write
    d.depth_compare_op = draw.depth_compare
    d.stencil_test_enable = draw.stencil_test
    d.front = _stencil_state(draw.front)
    d.back = _stencil_state(draw.back)
    b = abi.BlendState()
    b.blend_enable = draw.blend_enable
    (b.src_color_blend_factor, b.dst_color_blend_factor, b.color_blend_op,
     b.src_alpha_blend_factor, b.dst_alpha_blend_factor, b.alpha_blend_op) = draw.blend
    b.color_write_mask = draw.color_write_mask
    d.blend = b
    for k in range(4):
        d.blend_constants[k] = draw.blend_constants[k]
    d.shader = draw.shader
    if draw.texture is not None and buf.texture_levels:
        t = draw.texture
        s = abi.SamplerState()
        s.mag_filter, s.min_filter, s.mipmap_mode = t.mag_filter, t.min_filter, t.mipmap_mode
        s.address_mode_u, s.address_mode_v = t.address_u, t.address_v
        s.mip_lod_bias = t.mip_lod_bias
        s.anisotropy_enable = t.anisotropy_enable
        s.max_anisotropy = t.max_anisotropy
        s.min_lod, s.max_lod = t.min_lod, t.max_lod
        d.sampler = s
        tx = abi.Texture()
        tx.format = t.format
        levels = buf.texture_levels[:abi.MAX_MIP_LEVELS]
        tx.level_count = len(levels)
        for i, (addr, w, h, pitch) in enumerate(levels):
            tx.levels[i] = abi.ImageLevel(addr, w, h, pitch, 0)
        d.texture = tx
    d.framebuffer_width = draw.width
    d.framebuffer_height = draw.height
    if buf.color:
        d.color = abi.Attachment(buf.color, draw.color_format, buf.color_pitch, buf.color_sample_pitch)
    if buf.depth:
        d.depth = abi.Attachment(buf.depth, draw.depth_format, buf.depth_pitch, buf.depth_sample_pitch)
    if buf.stencil:
        d.stencil = abi.Attachment(buf.stencil, abi.FORMAT_S8_UINT, buf.stencil_pitch, buf.stencil_sample_pitch)
    d.occlusion_count = buf.occlusion or None
    d.topology = draw.topology
    d.provoking_vertex = draw.provoking_vertex
    if draw.depth_bias is not None:
        d.depth_bias_enable = 1
        (d.depth_bias_constant_factor, d.depth_bias_clamp, d.depth_bias_slope_factor) = draw.depth_bias
    d.line_width = draw.line_width
    d.line_rasterization_mode = draw.line_rasterization_mode
    return d
