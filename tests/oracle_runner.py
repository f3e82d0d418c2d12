"""Runs the CPU oracle (oracle/libswr_oracle.so) on a DrawData.

Test infrastructure: the oracle is the checker, never the product.
"""
import ctypes
import os
import subprocess

import numpy as np

from swiftshader_amd import abi
from swiftshader_amd.scenes import DrawData
from swiftshader_amd.desc import (COLOR_FORMATS, DEPTH_FORMATS, FORMAT_BYTES, Buffers, build_desc,
                                  color_bytes_per_pixel, depth_bytes_per_pixel)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "libswr_oracle.so")
REF_SO = os.path.join(ORACLE_DIR, "_ref", "libswr_ref.so")

_oracle = None
_ref = None


def build_oracle():
    src = os.path.join(ORACLE_DIR, "swr_oracle.c")
    if not os.path.exists(ORACLE_SO) or os.path.getmtime(ORACLE_SO) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR, "libswr_oracle.so"], check=True)


def oracle():
    global _oracle
    if _oracle is None:
        build_oracle()
        L = ctypes.CDLL(ORACLE_SO)
        L.swo_draw.argtypes = [ctypes.POINTER(abi.DrawDesc)]
        L.swo_draw.restype = ctypes.c_int
        L.swo_draw_mt.argtypes = [ctypes.POINTER(abi.DrawDesc), ctypes.c_int]
        L.swo_draw_mt.restype = ctypes.c_int
        L.swo_draw_state.argtypes = [ctypes.POINTER(abi.DrawDesc), ctypes.c_void_p, ctypes.c_int]
        L.swo_draw_state.restype = ctypes.c_int
        L.swo_sample_quad.argtypes = [ctypes.POINTER(abi.DrawDesc)] + [ctypes.c_void_p] * 4
        L.swo_sample_quad.restype = ctypes.c_int
        L.swo_resolve.argtypes = [ctypes.POINTER(abi.ResolveDesc)]
        L.swo_resolve.restype = ctypes.c_int
        L.swo_clear.argtypes = [ctypes.POINTER(abi.ClearDesc)]
        L.swo_clear.restype = ctypes.c_int
        for f in ("swo_round_int", "swo_round_int_clamped", "swo_trunc_int"):
            getattr(L, f).argtypes = [ctypes.c_float]
            getattr(L, f).restype = ctypes.c_int32
        for f in ("swo_mulhigh_u16x2", "swo_mulhigh_s16x2"):
            getattr(L, f).argtypes = [ctypes.c_uint32, ctypes.c_uint32]
            getattr(L, f).restype = ctypes.c_uint32
        for f in ("swo_log2sqrt", "swo_frac", "swo_linear_to_srgb", "swo_srgb_to_linear"):
            getattr(L, f).argtypes = [ctypes.c_float]
            getattr(L, f).restype = ctypes.c_float
        L.swo_constants.restype = None
        L.swo_srgb_lut.argtypes = [ctypes.c_void_p]
        L.swo_srgb_lut.restype = None
        L.swo_decode_bc.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_int] * 6
        L.swo_decode_bc.restype = ctypes.c_int
        L.swo_decode_etc.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_int] * 5
        L.swo_decode_etc.restype = ctypes.c_int
        fp = ctypes.POINTER(ctypes.c_float)
        L.swo_clip_triangle.argtypes = [fp, fp, fp, ctypes.c_int, fp, ctypes.POINTER(ctypes.c_int)]
        L.swo_clip_triangle.restype = ctypes.c_int
        _oracle = L
    return _oracle


def reference_lib():
    """oracle/_ref/libswr_ref.so (reference sources compiled as-is), or None."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        L = ctypes.CDLL(REF_SO)
        fp = ctypes.POINTER(ctypes.c_float)
        L.ref_clip_triangle.argtypes = [fp, fp, fp, ctypes.c_int, fp, ctypes.POINTER(ctypes.c_int)]
        L.ref_clip_triangle.restype = ctypes.c_int
        L.ref_constants.restype = None
        L.ref_srgb_lut.argtypes = [ctypes.c_void_p]
        L.ref_srgb_lut.restype = None
        L.ref_decode_bc.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_int] * 6
        L.ref_decode_bc.restype = ctypes.c_int
        L.ref_decode_etc.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_int] * 5
        L.ref_decode_etc.restype = ctypes.c_int
        _ref = L
    return _ref


def host_targets(draw):
    """Host attachments cleared by the oracle's clear (the same values the
    library's swr_clear writes)."""
    S, H, W = draw.samples, draw.height, draw.width
    if draw.color_format == abi.FORMAT_R32G32B32A32_SFLOAT:
        color = np.empty((S, H, W, 4), np.float32)
    else:
        color = np.empty((S, H, W, 4), np.uint8)
    oracle_clear(color, draw.color_format, draw.clear_color)
    depth = stencil = None
    if draw.has_depth:
        ddt = np.uint16 if draw.depth_format == abi.FORMAT_D16_UNORM else np.float32
        depth = oracle_clear(np.empty((S, H, W), ddt), draw.depth_format, draw.clear_depth)
    if draw.has_stencil:
        stencil = oracle_clear(np.empty((S, H, W), np.uint8), abi.FORMAT_S8_UINT, draw.clear_stencil)
    return color, depth, stencil


def host_threads():
    """Host threads for the multithreaded oracle: the CPUs this process may
    use, at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def oracle_render_ids(draw, threads=1):
    """Renders `draw` (RGBA32F target) with the primitive_id record; returns
    (color, depth, ids (S, H, W) uint32, occlusion)."""
    ids = np.zeros((draw.samples, draw.height, draw.width), np.uint32)
    color, depth, _, occ = oracle_render(draw, ids=ids, threads=threads)
    return color, depth, ids, occ


def oracle_render_state(draw, threads=1):
    """Renders `draw` (RGBA32F target) with the primitive_id record and the
    per-sample sampling state of the colour-writing fragment; returns
    (color, depth, ids (S, H, W) uint32, state (S, H, W, 4) float32
    (lod, anisotropy, uDelta, vDelta), occlusion)."""
    ids = np.zeros((draw.samples, draw.height, draw.width), np.uint32)
    state = np.zeros((draw.samples, draw.height, draw.width, 4), np.float32)
    color, depth, _, occ = oracle_render(draw, ids=ids, threads=threads, state=state)
    return color, depth, ids, state, occ


def _texture_desc(texture):
    """A draw descriptor holding only `texture` (for swo_sample_quad) and the arrays it points to."""
    keep = []
    levels = []
    for l in texture.levels:
        a = np.ascontiguousarray(l)
        keep.append(a)
        levels.append((a.ctypes.data, a.shape[1], a.shape[0], a.shape[1] * a.shape[2] * a.itemsize))
    pos = np.zeros((3, 4), np.float32)
    keep.append(pos)
    d = DrawData(width=8, height=8, positions=pos, texcoords=np.zeros((3, 2), np.float32), texture=texture,
                 shader=abi.SHADER_SAMPLE, color_format=abi.FORMAT_R32G32B32A32_SFLOAT, has_depth=False,
                 depth_test=0)
    buf = Buffers(positions=pos.ctypes.data, texcoords=pos.ctypes.data)
    buf.texture_levels = levels
    return build_desc(d, buf), keep


def oracle_sample_quad(texture, u, v):
    """sampleTexture128 of one quad: u, v are the 4 lanes (0,0), (1,0), (0,1),
    (1,1).  Returns (rgba (4, 4) float32, state (lod, anisotropy, uDelta,
    vDelta) float32); RGBA8 textures return the 16-bit result / 0xFF00."""
    L = oracle()
    desc, keep = _texture_desc(texture)
    uu = np.ascontiguousarray(u, np.float32)
    vv = np.ascontiguousarray(v, np.float32)
    out = np.zeros((4, 4), np.float32)
    st = np.zeros(4, np.float32)
    rc = L.swo_sample_quad(ctypes.byref(desc), uu.ctypes.data, vv.ctypes.data, out.ctypes.data, st.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"oracle swo_sample_quad failed: {rc}")
    return out, st


def oracle_render(draw, targets=None, ids=None, threads=1, state=None):
    """Renders `draw` with the oracle; returns (color, depth, stencil, occlusion).

    threads > 1 runs swo_draw_mt (row pairs split over host threads, the
    reference's cluster split), which renders the same bytes.  `state`, an
    (S, H, W, 4) float32 array, receives the sampling state of each sample's
    colour-writing fragment."""
    L = oracle()
    color, depth, stencil = targets if targets is not None else host_targets(draw)
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a.ctypes.data

    buf = Buffers(positions=arr(draw.positions, np.float32))
    if draw.colors is not None:
        buf.colors = arr(draw.colors, np.float32)
    if draw.texcoords is not None:
        buf.texcoords = arr(draw.texcoords, np.float32)
    if draw.indices is not None:
        ind = draw.indices
        buf.indices = arr(ind, ind.dtype if ind.dtype in (np.uint8, np.uint16) else np.uint32)
    if draw.point_sizes is not None:
        buf.point_sizes = arr(draw.point_sizes, np.float32)
    if draw.texture is not None:
        levels = []
        for l in draw.texture.levels:
            a = np.ascontiguousarray(l)
            keep.append(a)
            levels.append((a.ctypes.data, a.shape[1], a.shape[0], a.shape[1] * a.shape[2] * a.itemsize))
        buf.texture_levels = levels
    S, H, W = draw.samples, draw.height, draw.width
    bpp = color_bytes_per_pixel(draw.color_format)
    buf.color, buf.color_pitch, buf.color_sample_pitch = color.ctypes.data, W * bpp, H * W * bpp
    if depth is not None:
        dbpp = depth_bytes_per_pixel(draw.depth_format)
        buf.depth, buf.depth_pitch, buf.depth_sample_pitch = depth.ctypes.data, W * dbpp, H * W * dbpp
    if stencil is not None:
        buf.stencil, buf.stencil_pitch, buf.stencil_sample_pitch = stencil.ctypes.data, W, H * W
    occ = np.zeros(1, np.uint64)
    buf.occlusion = occ.ctypes.data
    desc = build_desc(draw, buf)
    if ids is not None:
        desc.primitive_id = abi.Attachment(ids.ctypes.data, abi.FORMAT_R32_UINT, W * 4, H * W * 4)
    if state is not None:
        assert state.shape == (S, H, W, 4) and state.dtype == np.float32 and state.flags["C_CONTIGUOUS"]
        rc = L.swo_draw_state(ctypes.byref(desc), state.ctypes.data, int(threads))
    elif threads > 1:
        rc = L.swo_draw_mt(ctypes.byref(desc), int(threads))
    else:
        rc = L.swo_draw(ctypes.byref(desc))
    if rc != 0:
        raise RuntimeError(f"oracle swo_draw failed: {rc}")
    return color, depth, stencil, int(occ[0])


def oracle_resolve(src, fmt=None):
    """4x -> 1x resolve of a (4, H, W, 4) array with the oracle (format by
    dtype unless given)."""
    L = oracle()
    S, H, W, _ = src.shape
    src = np.ascontiguousarray(src)
    bpp = 16 if src.dtype == np.float32 else 4
    if fmt is None:
        fmt = abi.FORMAT_R32G32B32A32_SFLOAT if src.dtype == np.float32 else abi.FORMAT_R8G8B8A8_UNORM
    dst = np.zeros((1, H, W, 4), src.dtype)
    r = abi.ResolveDesc(W, H, fmt, src.ctypes.data, W * bpp, H * W * bpp, S, dst.ctypes.data, W * bpp)
    rc = L.swo_resolve(ctypes.byref(r))
    if rc != 0:
        raise RuntimeError(f"oracle swo_resolve failed: {rc}")
    return dst


def oracle_clear(target, fmt, value, area=None):
    """Clears a (S, H, W[, C]) host array in place with the oracle (same
    descriptor as the library's swr_clear)."""
    L = oracle()
    assert target.flags["C_CONTIGUOUS"]
    S, H, W = target.shape[:3]
    bpp = FORMAT_BYTES[fmt]
    d = abi.ClearDesc()
    d.target = abi.Attachment(target.ctypes.data, fmt, W * bpp, H * W * bpp)
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
    rc = L.swo_clear(ctypes.byref(d))
    if rc != 0:
        raise RuntimeError(f"oracle swo_clear failed: {rc}")
    return target


def bc_block_bytes(n):
    return 8 if n in (1, 4) else 16


def oracle_decode_bc(src, dst, w, h, pitch, bpp, n, no_alpha_or_unsigned, lib=None):
    """Decodes BCn blocks `src` (uint8) into `dst` (uint8, pitch bytes per
    row) in place with the oracle, or with `lib`'s ref_decode_bc."""
    src = np.ascontiguousarray(src, np.uint8)
    assert dst.flags["C_CONTIGUOUS"] and dst.dtype == np.uint8
    if lib is None:
        return oracle().swo_decode_bc(src.ctypes.data, dst.ctypes.data, w, h, pitch, bpp, n, no_alpha_or_unsigned)
    return lib.ref_decode_bc(src.ctypes.data, dst.ctypes.data, w, h, pitch, bpp, n, no_alpha_or_unsigned)


ETC_TYPES = {"r11s": 0, "r11": 1, "rg11s": 2, "rg11": 3, "rgb8": 4, "rgb8a1": 5, "rgba8": 6}


def etc_block_bytes(t):
    return 16 if t in (2, 3, 6) else 8


def etc_out_bpp(t):
    return 2 if t in (0, 1) else 4


def oracle_decode_etc(src, dst, w, h, pitch, bpp, t, lib=None):
    """Decodes ETC2 / EAC blocks `src` into `dst` (uint8, pitch bytes per
    row) in place with the oracle, or with `lib`'s ref_decode_etc."""
    src = np.ascontiguousarray(src, np.uint8)
    assert dst.flags["C_CONTIGUOUS"] and dst.dtype == np.uint8
    if lib is None:
        return oracle().swo_decode_etc(src.ctypes.data, dst.ctypes.data, w, h, pitch, bpp, t)
    return lib.ref_decode_etc(src.ctypes.data, dst.ctypes.data, w, h, pitch, bpp, t)
