"""Differentiable draws: torch.autograd.Function over the C ABI.

Forward passes run the draw path (``swr_draw``) and the box-filtered mip
chain (``swr_generate_mips``); backward passes run the hand-written HIP
kernels behind ``swr_draw_backward`` and ``swr_mip_backward``.  PyTorch only
holds the tensors and orders the calls on its current stream.

What is differentiated (the gradient has no counterpart in the reference):
coverage, visibility, the texture LOD and the anisotropic probe layout are
discrete and held at the forward pass's values; each covered sample takes
its colour from the primitive recorded in the draw's primitive_id
attachment, interpolated at the pixel centre with perspective-correct
barycentrics of the unsnapped projected vertices and filtered as the forward
filters.  Gradients reach texels (through the filter footprint and the mip
chain), per-vertex colours and texcoords, and vertex x, y, w (z only sets
visibility, so its gradient is zero).
"""
import ctypes
from dataclasses import replace
from typing import List, Optional, Sequence

import torch

from . import abi, runtime
from .desc import Buffers, build_desc
from .scenes import DrawData, TextureData


def _texture_struct(levels: Sequence[torch.Tensor], fmt: int) -> abi.Texture:
    t = abi.Texture()
    t.format = fmt
    t.level_count = len(levels)
    for i, l in enumerate(levels):
        t.levels[i] = abi.ImageLevel(l.data_ptr(), l.shape[1], l.shape[0], l.shape[1] * l.shape[2] * l.element_size(), 0)
    return t


def mip_levels_like(level0: torch.Tensor, count: Optional[int] = None) -> List[torch.Tensor]:
    """Empty tensors for levels 1.. of the box chain of `level0` (H, W, 4)."""
    out = []
    h, w = level0.shape[:2]
    while (h > 1 or w > 1) and (count is None or len(out) + 1 < count):
        h, w = max(h // 2, 1), max(w // 2, 1)
        out.append(torch.empty((h, w, 4), dtype=level0.dtype, device=level0.device))
    return out


def generate_mips(levels: Sequence[torch.Tensor]):
    """Fills levels[1:] from levels[0] (swr_generate_mips) on the current stream."""
    fmt = abi.FORMAT_R32G32B32A32_SFLOAT if levels[0].dtype == torch.float32 else abi.FORMAT_R8G8B8A8_UNORM
    t = _texture_struct(levels, fmt)
    runtime.check(runtime.lib().swr_generate_mips(ctypes.byref(t), runtime._stream_handle(None)), "swr_generate_mips")


class MipChain(torch.autograd.Function):
    """level0 (H, W, 4) float32 -> levels 1.. of its 2x2 box chain."""

    @staticmethod
    def forward(ctx, level0: torch.Tensor, count: int):
        rest = mip_levels_like(level0, count)
        generate_mips([level0] + rest)
        ctx.shapes = [level0.shape] + [r.shape for r in rest]
        return tuple(rest)

    @staticmethod
    def backward(ctx, *grads):
        dev = grads[0].device if grads and grads[0] is not None else None
        bufs = [torch.zeros(ctx.shapes[0], dtype=torch.float32, device=dev)]
        for shp, g in zip(ctx.shapes[1:], grads):
            bufs.append(g.contiguous().clone() if g is not None else torch.zeros(shp, dtype=torch.float32, device=dev))
        t = _texture_struct(bufs, abi.FORMAT_R32G32B32A32_SFLOAT)
        runtime.check(runtime.lib().swr_mip_backward(ctypes.byref(t), runtime._stream_handle(None)),
                      "swr_mip_backward")
        return bufs[0], None


def texture_chain(level0: torch.Tensor, count: Optional[int] = None) -> List[torch.Tensor]:
    """[level0, level1, ...], differentiable with respect to level0."""
    h, w = level0.shape[:2]
    full = 1
    while h > 1 or w > 1:
        h, w = max(h // 2, 1), max(w // 2, 1)
        full += 1
    count = full if count is None else min(count, full)
    if count <= 1:
        return [level0]
    return [level0] + list(MipChain.apply(level0, count))


class View:
    """The fixed part of one differentiable draw: state, sampler, viewport,
    and device buffers for depth and the primitive_id record.  The colour
    attachment is RGBA32F; blending must be off."""

    def __init__(self, draw: DrawData, device="cuda:0"):
        if draw.color_format != abi.FORMAT_R32G32B32A32_SFLOAT:
            raise ValueError("differentiable draws render into RGBA32F")
        if draw.blend_enable:
            raise ValueError("differentiable draws do not blend")
        if draw.depth_format != abi.FORMAT_D32_SFLOAT:
            raise ValueError("differentiable draws use a D32F depth attachment")
        self.draw = draw
        self.device = torch.device(device)
        S, H, W = draw.samples, draw.height, draw.width
        self.shape = (S, H, W)
        self.depth = torch.empty((S, H, W), dtype=torch.float32, device=self.device) if draw.has_depth else None
        self.ids = torch.empty((S, H, W), dtype=torch.int32, device=self.device)
        self.indices = None
        if draw.indices is not None:
            self.indices = torch.from_numpy(draw.indices.astype("uint32").view("int32")).to(self.device)

    def desc(self, pos, col, uv, levels, color) -> abi.DrawDesc:
        S, H, W = self.shape
        d = self.draw
        buf = Buffers(positions=pos.data_ptr(), colors=col.data_ptr() if col is not None else 0,
                      texcoords=uv.data_ptr() if uv is not None else 0,
                      indices=self.indices.data_ptr() if self.indices is not None else 0)
        if levels:
            buf.texture_levels = [(l.data_ptr(), l.shape[1], l.shape[0], l.shape[1] * 16) for l in levels]
        buf.color, buf.color_pitch, buf.color_sample_pitch = color.data_ptr(), W * 16, H * W * 16
        if self.depth is not None:
            buf.depth, buf.depth_pitch, buf.depth_sample_pitch = self.depth.data_ptr(), W * 4, H * W * 4
        # build_desc needs a TextureData for the sampler state; the level
        # addresses come from `buf`.
        tex = d.texture
        if levels and tex is None:
            tex = TextureData(format=abi.FORMAT_R32G32B32A32_SFLOAT, levels=[])
        dd = build_desc(replace(d, texture=tex) if tex is not d.texture else d, buf)
        dd.vertex_count = pos.shape[0]
        dd.index_type = abi.INDEX_TYPE_UINT32  # the view uploads its indices as uint32
        dd.primitive_id = abi.Attachment(self.ids.data_ptr(), abi.FORMAT_R32_UINT, W * 4, H * W * 4)
        return dd

    def clear(self, color: torch.Tensor):
        d = self.draw
        color[:] = torch.tensor(d.clear_color, dtype=torch.float32, device=self.device)
        if self.depth is not None:
            self.depth.fill_(float(d.clear_depth))
        self.ids.zero_()


class Draw(torch.autograd.Function):
    """One differentiable draw: (positions, colours, texcoords, texture
    levels) -> colour attachment (S, H, W, 4) float32."""

    @staticmethod
    def forward(ctx, view: View, pos, col, uv, *levels):
        pos, col, uv = [t.contiguous() if t is not None else None for t in (pos, col, uv)]
        levels = [l.contiguous() for l in levels]
        color = torch.empty(view.shape + (4,), dtype=torch.float32, device=view.device)
        view.clear(color)
        desc = view.desc(pos, col, uv, levels, color)
        runtime.check(runtime.lib().swr_draw(ctypes.byref(desc), runtime._stream_handle(None)), "swr_draw")
        ctx.view = view
        ctx.has = (col is not None, uv is not None)
        ctx.save_for_backward(pos, col if col is not None else pos, uv if uv is not None else pos, *levels)
        ctx.ids = view.ids.clone()
        return color

    @staticmethod
    def backward(ctx, d_color):
        view = ctx.view
        pos, col, uv, *levels = ctx.saved_tensors
        col = col if ctx.has[0] else None
        uv = uv if ctx.has[1] else None
        d_color = d_color.contiguous()
        S, H, W = view.shape
        d_pos = torch.zeros_like(pos)
        d_col = torch.zeros_like(col) if col is not None else None
        d_uv = torch.zeros_like(uv) if uv is not None else None
        d_lv = [torch.zeros_like(l) for l in levels]
        # The visibility record of this forward call (the view's buffer may
        # have been overwritten by a later forward).
        saved_ids = view.ids
        view.ids = ctx.ids
        color = torch.empty(view.shape + (4,), dtype=torch.float32, device=view.device)
        desc = view.desc(pos, col, uv, levels, color)
        view.ids = saved_ids
        g = abi.GradDesc()
        g.d_color = d_color.data_ptr()
        g.d_color_row_pitch_bytes = W * 16
        g.d_color_sample_pitch_bytes = H * W * 16
        g.d_positions = d_pos.data_ptr()
        g.d_colors = d_col.data_ptr() if d_col is not None else None
        g.d_texcoords = d_uv.data_ptr() if d_uv is not None else None
        for i, t in enumerate(d_lv):
            g.d_texture[i] = t.data_ptr()
        runtime.check(runtime.lib().swr_draw_backward(ctypes.byref(desc), ctypes.byref(g),
                                                      runtime._stream_handle(None)), "swr_draw_backward")
        return (None, d_pos, d_col, d_uv, *d_lv)


def draw(view: View, positions, colors=None, texcoords=None, levels: Sequence[torch.Tensor] = ()):
    """Differentiable draw; returns the (S, H, W, 4) float32 colour attachment."""
    return Draw.apply(view, positions, colors, texcoords, *levels)


def backward_raw(view: View, positions, colors, texcoords, levels, ids, d_color, sample_state=None):
    """Direct swr_draw_backward call (no autograd): returns the gradient
    buffers, and fills `sample_state` (S, H, W, 4) if given."""
    S, H, W = view.shape
    d_pos = torch.zeros_like(positions)
    d_col = torch.zeros_like(colors) if colors is not None else None
    d_uv = torch.zeros_like(texcoords) if texcoords is not None else None
    d_lv = [torch.zeros_like(l) for l in levels]
    saved = view.ids
    view.ids = ids
    color = torch.empty(view.shape + (4,), dtype=torch.float32, device=view.device)
    desc = view.desc(positions, colors, texcoords, list(levels), color)
    view.ids = saved
    g = abi.GradDesc()
    g.d_color = d_color.data_ptr()
    g.d_color_row_pitch_bytes = W * 16
    g.d_color_sample_pitch_bytes = H * W * 16
    g.d_positions = d_pos.data_ptr()
    g.d_colors = d_col.data_ptr() if d_col is not None else None
    g.d_texcoords = d_uv.data_ptr() if d_uv is not None else None
    for i, t in enumerate(d_lv):
        g.d_texture[i] = t.data_ptr()
    if sample_state is not None:
        g.sample_state = sample_state.data_ptr()
        g.sample_state_row_pitch_bytes = W * 16
        g.sample_state_sample_pitch_bytes = H * W * 16
    runtime.check(runtime.lib().swr_draw_backward(ctypes.byref(desc), ctypes.byref(g), runtime._stream_handle(None)),
                  "swr_draw_backward")
    return d_pos, d_col, d_uv, d_lv
