"""Python host side of the draw path: device buffers, clears and draws.

PyTorch provides device memory and streams; all rendering work runs in the
HIP kernels of ``libswr.so`` through the C ABI of ``include/swr.h``.  There
is no fallback: if the shared library is missing or fails to load, using
this module raises.
"""
import ctypes
import os
from dataclasses import dataclass
from typing import List, Optional

import numpy as np
import torch

from . import abi
from .desc import COLOR_FORMATS, DEPTH_FORMATS, FORMAT_BYTES, Buffers, build_desc, color_bytes_per_pixel, \
    depth_bytes_per_pixel
from .scenes import DrawData

_LIB = None
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libswr.so")


class SwrError(RuntimeError):
    pass


def lib():
    """The loaded libswr (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise SwrError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        _LIB = abi.bind(ctypes.CDLL(LIB_PATH))
        if _LIB.swr_abi_version() != abi.ABI_VERSION:
            raise SwrError("libswr ABI version mismatch")
    return _LIB


def check(rc: int, what: str = "swr call"):
    if rc != abi.OK:
        raise SwrError(f"{what} failed ({rc}): {lib().swr_last_error().decode()}")


def _stream_handle(stream) -> Optional[int]:
    if stream is None:
        stream = torch.cuda.current_stream()
    return stream.cuda_stream


@dataclass
class Targets:
    """Render targets of one draw: sample-major planes, row-major pixels."""
    color: Optional[torch.Tensor]    # (S, H, W, 4) uint8 or float32
    depth: Optional[torch.Tensor]    # (S, H, W) float32 (D32F) or int16 holding D16's bits
    stencil: Optional[torch.Tensor]  # (S, H, W) uint8
    occlusion: torch.Tensor          # (1,) int64


def clear_targets(draw: DrawData, targets: Targets, stream=None, reset_occlusion: bool = True):
    """Clears the draw's attachments to its clear values (render-pass load op
    CLEAR) with the library's swr_clear; resets the occlusion counter unless
    told not to."""
    if targets.color is not None:
        clear(targets.color, draw.color_format, draw.clear_color, stream=stream)
    if targets.depth is not None:
        clear(targets.depth, draw.depth_format, draw.clear_depth, stream=stream)
    if targets.stencil is not None:
        clear(targets.stencil, abi.FORMAT_S8_UINT, draw.clear_stencil, stream=stream)
    if reset_occlusion:
        targets.occlusion.zero_()


def make_targets(draw: DrawData, device) -> Targets:
    S, H, W = draw.samples, draw.height, draw.width
    if draw.color_format == abi.FORMAT_R32G32B32A32_SFLOAT:
        color = torch.empty((S, H, W, 4), dtype=torch.float32, device=device)
    else:
        color = torch.empty((S, H, W, 4), dtype=torch.uint8, device=device)
    ddt = torch.int16 if draw.depth_format == abi.FORMAT_D16_UNORM else torch.float32  # D16: 16-bit storage
    depth = torch.empty((S, H, W), dtype=ddt, device=device) if draw.has_depth else None
    stencil = torch.empty((S, H, W), dtype=torch.uint8, device=device) if draw.has_stencil else None
    occ = torch.zeros(1, dtype=torch.int64, device=device)
    t = Targets(color, depth, stencil, occ)
    with torch.cuda.device(torch.device(device)):
        clear_targets(draw, t)
    return t


@dataclass
class DeviceDraw:
    """A draw whose inputs live on the device, with its C descriptor."""
    draw: DrawData
    tensors: list
    targets: Targets
    desc: abi.DrawDesc


def _dev(a: np.ndarray, device) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def upload_texture(tex, device) -> List[torch.Tensor]:
    return [_dev(l, device) for l in tex.levels]


def prepare(draw: DrawData, device, targets: Optional[Targets] = None,
            texture_tensors: Optional[List[torch.Tensor]] = None) -> DeviceDraw:
    keep = []
    pos = _dev(draw.positions.astype(np.float32), device)
    keep.append(pos)
    buf = Buffers(positions=pos.data_ptr())
    if draw.colors is not None:
        col = _dev(draw.colors.astype(np.float32), device)
        keep.append(col)
        buf.colors = col.data_ptr()
    if draw.texcoords is not None:
        uv = _dev(draw.texcoords.astype(np.float32), device)
        keep.append(uv)
        buf.texcoords = uv.data_ptr()
    if draw.indices is not None:
        ind = np.ascontiguousarray(draw.indices)
        if ind.dtype == np.uint16:
            idx = _dev(ind.view(np.int16), device)
        elif ind.dtype == np.uint8:
            idx = _dev(ind, device)
        else:
            idx = _dev(ind.astype(np.uint32).view(np.int32), device)
        keep.append(idx)
        buf.indices = idx.data_ptr()
    if draw.point_sizes is not None:
        ps = _dev(draw.point_sizes.astype(np.float32), device)
        keep.append(ps)
        buf.point_sizes = ps.data_ptr()
    if draw.texture is not None:
        levels = texture_tensors if texture_tensors is not None else upload_texture(draw.texture, device)
        keep.extend(levels)
        buf.texture_levels = [(t.data_ptr(), t.shape[1], t.shape[0], t.shape[1] * t.shape[2] * t.element_size())
                              for t in levels]
    if targets is None:
        targets = make_targets(draw, device)
    S, H, W = draw.samples, draw.height, draw.width
    bpp = color_bytes_per_pixel(draw.color_format)
    if targets.color is not None:
        buf.color, buf.color_pitch, buf.color_sample_pitch = targets.color.data_ptr(), W * bpp, H * W * bpp
    if targets.depth is not None:
        dbpp = depth_bytes_per_pixel(draw.depth_format)
        buf.depth, buf.depth_pitch, buf.depth_sample_pitch = targets.depth.data_ptr(), W * dbpp, H * W * dbpp
    if targets.stencil is not None:
        buf.stencil, buf.stencil_pitch, buf.stencil_sample_pitch = targets.stencil.data_ptr(), W, H * W
    buf.occlusion = targets.occlusion.data_ptr()
    return DeviceDraw(draw, keep, targets, build_desc(draw, buf))


def draw(dd: DeviceDraw, stream=None):
    """Enqueues one draw (sw::Renderer::draw) on `stream`."""
    check(lib().swr_draw(ctypes.byref(dd.desc), _stream_handle(stream)), "swr_draw")


def draw_batch(dds: List[DeviceDraw], stream=None):
    arr = (abi.DrawDesc * len(dds))(*[d.desc for d in dds])
    check(lib().swr_draw_batch(arr, len(dds), _stream_handle(stream)), "swr_draw_batch")


def resolve(src: torch.Tensor, dst: torch.Tensor, stream=None):
    """4x -> 1x colour resolve (sw::Blitter::resolve); src (4,H,W,4), dst (1,H,W,4).

    Pixels and channels must be contiguous; rows and sample planes may lie
    apart (the pitches come from the strides)."""
    S, H, W, _ = src.shape
    fmt = abi.FORMAT_R32G32B32A32_SFLOAT if src.dtype == torch.float32 else abi.FORMAT_R8G8B8A8_UNORM
    bpp = 16 if src.dtype == torch.float32 else 4
    es = src.element_size()
    for t in (src, dst):
        if t.stride(2) * t.element_size() != bpp or t.stride(3) != 1:
            raise ValueError("resolve needs contiguous pixels (H, W, 4) within each row")
    r = abi.ResolveDesc(W, H, fmt, src.data_ptr(), src.stride(1) * es, src.stride(0) * es, S, dst.data_ptr(),
                        dst.stride(1) * dst.element_size())
    check(lib().swr_resolve(ctypes.byref(r), _stream_handle(stream)), "swr_resolve")


def clear_batch(items, stream=None):
    """Clears several targets with one swr_clear_batch call: items is a list
    of (target, fmt, value) or (target, fmt, value, area), as clear() takes
    them; the targets must not overlap."""
    descs = (abi.ClearDesc * max(1, len(items)))()
    for i, it in enumerate(items):
        descs[i] = clear_desc(*it)
    check(lib().swr_clear_batch(descs, len(items), _stream_handle(stream)), "swr_clear_batch")


def clear_targets_batch(pairs, stream=None):
    """clear_targets for many (draw, targets) pairs in one swr_clear_batch call
    (occlusion counters untouched)."""
    items = []
    for draw, t in pairs:
        if t.color is not None:
            items.append((t.color, draw.color_format, draw.clear_color))
        if t.depth is not None:
            items.append((t.depth, draw.depth_format, draw.clear_depth))
        if t.stencil is not None:
            items.append((t.stencil, abi.FORMAT_S8_UINT, draw.clear_stencil))
    clear_batch(items, stream)


def clear_desc(target: torch.Tensor, fmt: int, value, area=None) -> abi.ClearDesc:
    """A swr_clear_desc for a (S, H, W[, C]) target tensor of format fmt.

    value: 4 floats (colour), one float (D32F or D16 depth) or one int (S8
    stencil); area: (x, y, width, height) or None for the whole image."""
    S, H, W = target.shape[:3]
    bpp = FORMAT_BYTES[fmt]
    es = target.element_size()
    if target.stride(2) * es != bpp:
        raise ValueError("clear needs contiguous pixels within each row")
    d = abi.ClearDesc()
    d.target = abi.Attachment(target.data_ptr(), fmt, target.stride(1) * es, target.stride(0) * es)
    d.width, d.height, d.samples = W, H, S
    if area is not None:
        d.use_area = 1
        d.area = abi.Rect(*area)
    if fmt in COLOR_FORMATS:
        for k in range(4):
            d.value.color[k] = float(value[k])
    elif fmt in DEPTH_FORMATS:
        d.value.depth_stencil.depth = float(value)
    else:
        d.value.depth_stencil.stencil = int(value) & 0xFFFFFFFF
    return d


def clear(target: torch.Tensor, fmt: int, value, area=None, stream=None):
    """Attachment clear (vk::Image::clear -> sw::Blitter::clear) of every sample of `area`."""
    d = clear_desc(target, fmt, value, area)
    check(lib().swr_clear(ctypes.byref(d), _stream_handle(stream)), "swr_clear")


def profile_enable(enable: bool = True):
    check(lib().swr_profile_enable(1 if enable else 0), "swr_profile_enable")


def profile_read() -> dict:
    """Per-stage GPU milliseconds summed over the draws since the last read."""
    p = abi.Profile()
    check(lib().swr_profile_read(ctypes.byref(p)), "swr_profile_read")
    return {"setup_ms": p.setup_ms, "bin_ms": p.bin_ms, "raster_ms": p.raster_ms, "total_ms": p.total_ms,
            "draws": int(p.draws), "pairs": int(p.pairs), "block_tests": int(p.block_tests),
            "block_covers": int(p.block_covers), "shaded": int(p.shaded)}


def render(draw_data: DrawData, device="cuda:0", stream=None):
    """Convenience: upload, clear, draw, return the targets (device tensors)."""
    dd = prepare(draw_data, device)
    draw(dd, stream)
    return dd.targets
