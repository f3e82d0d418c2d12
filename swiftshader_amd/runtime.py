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
from .desc import Buffers, build_desc, color_bytes_per_pixel
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
    depth: Optional[torch.Tensor]    # (S, H, W) float32
    stencil: Optional[torch.Tensor]  # (S, H, W) uint8
    occlusion: torch.Tensor          # (1,) int64


def clear_value_rgba8(c):
    return [int(np.clip(np.rint(np.float32(v) * np.float32(255.0)), 0, 255)) for v in c]


def make_targets(draw: DrawData, device) -> Targets:
    S, H, W = draw.samples, draw.height, draw.width
    if draw.color_format == abi.FORMAT_R32G32B32A32_SFLOAT:
        color = torch.empty((S, H, W, 4), dtype=torch.float32, device=device)
        color[:] = torch.tensor(draw.clear_color, dtype=torch.float32, device=device)
    else:
        color = torch.empty((S, H, W, 4), dtype=torch.uint8, device=device)
        color[:] = torch.tensor(clear_value_rgba8(draw.clear_color), dtype=torch.uint8, device=device)
    depth = None
    if draw.has_depth:
        depth = torch.full((S, H, W), float(draw.clear_depth), dtype=torch.float32, device=device)
    stencil = None
    if draw.has_stencil:
        stencil = torch.full((S, H, W), int(draw.clear_stencil), dtype=torch.uint8, device=device)
    occ = torch.zeros(1, dtype=torch.int64, device=device)
    return Targets(color, depth, stencil, occ)


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
        idx = _dev(draw.indices.astype(np.uint32).view(np.int32), device)
        keep.append(idx)
        buf.indices = idx.data_ptr()
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
        buf.depth, buf.depth_pitch, buf.depth_sample_pitch = targets.depth.data_ptr(), W * 4, H * W * 4
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
    """4x -> 1x colour resolve (sw::Blitter::resolve); src (4,H,W,4), dst (1,H,W,4)."""
    S, H, W, _ = src.shape
    fmt = abi.FORMAT_R32G32B32A32_SFLOAT if src.dtype == torch.float32 else abi.FORMAT_R8G8B8A8_UNORM
    bpp = 16 if src.dtype == torch.float32 else 4
    r = abi.ResolveDesc(W, H, fmt, src.data_ptr(), W * bpp, H * W * bpp, S, dst.data_ptr(), W * bpp)
    check(lib().swr_resolve(ctypes.byref(r), _stream_handle(stream)), "swr_resolve")


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
