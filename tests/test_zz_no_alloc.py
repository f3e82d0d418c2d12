"""swr_draw_noalloc, the draw whose scratch is a fixed device array, against
swr_draw, the entry point it is a variant of: the same images, depth,
stencil and occlusion counts, bit for bit.

* two calls in a row on one stream, the second with a smaller draw (the
  array keeps the first call's contents);
* a 4x MSAA draw with stencil;
* a draw whose (tile, primitive) pairs exceed the room the array leaves, so
  that the primitives past the capacity go through the unbinned pass;
* a draw too large for the array, refused with an error before any work.
"""
import ctypes
import dataclasses

import numpy as np
import pytest
import torch

from swiftshader_amd import abi, runtime, scenes

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _noalloc(dd):
    return runtime.lib().swr_draw_noalloc(ctypes.byref(dd.desc), runtime._stream_handle(None))


def _reference(draw):
    dd = runtime.prepare(draw, DEV)
    runtime.draw(dd)
    return dd


def _assert_same(a, b, what):
    ta, tb = a.targets, b.targets
    assert torch.equal(ta.color, tb.color), f"{what}: colour differs"
    if ta.depth is not None:
        assert torch.equal(ta.depth.view(torch.int32) if ta.depth.dtype == torch.float32 else ta.depth,
                           tb.depth.view(torch.int32) if tb.depth.dtype == torch.float32 else tb.depth), \
            f"{what}: depth differs"
    if ta.stencil is not None:
        assert torch.equal(ta.stencil, tb.stencil), f"{what}: stencil differs"
    assert int(ta.occlusion.item()) == int(tb.occlusion.item()), f"{what}: occlusion differs"


def _stencil_msaa(seed):
    d = scenes.plain_gouraud(seed=seed, triangles=3000, size=320)
    return dataclasses.replace(d, samples=4, has_stencil=True, stencil_test=1,
                               front=dict(fail_op=abi.STENCIL_OP_KEEP, pass_op=abi.STENCIL_OP_INCREMENT_AND_WRAP,
                                          depth_fail_op=abi.STENCIL_OP_DECREMENT_AND_CLAMP,
                                          compare_op=abi.COMPARE_ALWAYS, compare_mask=0xFF, write_mask=0xFF,
                                          reference=0),
                               back=dict(fail_op=abi.STENCIL_OP_KEEP, pass_op=abi.STENCIL_OP_INVERT,
                                         depth_fail_op=abi.STENCIL_OP_KEEP, compare_op=abi.COMPARE_ALWAYS,
                                         compare_mask=0xFF, write_mask=0xFF, reference=0))


def test_two_calls_in_a_row_second_smaller():
    big = scenes.textured_1080p(seed=11, triangles=30000, width=960, height=540, tex_size=512)
    small = scenes.plain_gouraud(seed=12, triangles=500, size=128)
    ref_big, ref_small = _reference(big), _reference(small)
    a, b = runtime.prepare(big, DEV), runtime.prepare(small, DEV)
    torch.cuda.synchronize()
    print("swr_draw_noalloc: 960x540 30k textured, then 128x128 500 gouraud", flush=True)
    runtime.check(_noalloc(a), "swr_draw_noalloc (first)")
    runtime.check(_noalloc(b), "swr_draw_noalloc (second, smaller)")
    torch.cuda.synchronize()
    _assert_same(ref_big, a, "first call")
    _assert_same(ref_small, b, "second call")


def test_msaa_stencil_matches_swr_draw():
    d = _stencil_msaa(13)
    ref = _reference(d)
    dd = runtime.prepare(d, DEV)
    print("swr_draw_noalloc: 320x320 4x MSAA stencil, 3000 triangles", flush=True)
    runtime.check(_noalloc(dd), "swr_draw_noalloc")
    torch.cuda.synchronize()
    _assert_same(ref, dd, "4x MSAA stencil")


def test_pairs_past_the_capacity_are_rendered():
    # 150k triangles of the benchmark's distribution at 1920x1080 need about
    # 4.4 M pairs; the array leaves room for about 3.4 M after their records.
    d = scenes.textured_1080p(seed=14, triangles=150_000, tex_size=256)
    ref = _reference(d)
    dd = runtime.prepare(d, DEV)
    print("swr_draw_noalloc: 1920x1080 150k textured (pairs past the capacity)", flush=True)
    runtime.check(_noalloc(dd), "swr_draw_noalloc")
    torch.cuda.synchronize()
    _assert_same(ref, dd, "overflowing draw")


def test_too_large_a_draw_is_refused():
    d = scenes.plain_gouraud(seed=15, triangles=400_000, size=256)
    dd = runtime.prepare(d, DEV)
    torch.cuda.synchronize()
    before = dd.targets.color.clone()
    rc = _noalloc(dd)
    torch.cuda.synchronize()
    assert rc == abi.ERROR_UNSUPPORTED
    assert b"scratch" in runtime.lib().swr_last_error()
    assert torch.equal(before, dd.targets.color)
