"""Batched draws on the GPU.

* swr_draw_batch of heterogeneous draws (every raster variant, and draws
  that share a target, so the batch splits into ordered segments) equals
  the oracle rendering the same draws one after another, bit for bit;
* a batch plan's runs equal the eager batch, and stay equal when the plan
  is captured into a HIP graph and replayed several times (with other
  device work between replays);
* every candidate tuning of the autotuner renders the same bytes.
"""
import ctypes

import numpy as np
import pytest
import torch

from cases import parity_cases
from oracle_runner import host_targets, oracle_render
from swiftshader_amd import abi, runtime, scenes

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _snap(dds):
    out = []
    for dd in dds:
        t = dd.targets
        out.append((t.color.cpu().numpy().copy(), None if t.depth is None else t.depth.cpu().numpy().copy(),
                    None if t.stencil is None else t.stencil.cpu().numpy().copy(), int(t.occlusion.item())))
    return out


def _reset(dds):
    for dd in dds:
        fresh = runtime.make_targets(dd.draw, DEV)
        t = dd.targets
        t.color.copy_(fresh.color)
        if t.depth is not None:
            t.depth.copy_(fresh.depth)
        if t.stencil is not None:
            t.stencil.copy_(fresh.stencil)
        t.occlusion.zero_()


def _same(a, b, name):
    for k, (x, y) in enumerate(zip(a, b)):
        if x is None:
            assert y is None
            continue
        if isinstance(x, int):
            assert x == y, (name, "occlusion")
            continue
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), (name, k)


def mixed_batch():
    """All parity scenes, plus two draws into one shared target."""
    cases = parity_cases()
    draws = [d for _, d in cases]
    dds = [runtime.prepare(d, DEV) for d in draws]
    rng = np.random.default_rng(9)
    pos, uv, col = scenes.screen_triangles(rng, 400, 64, 64, area_hi=600)
    a = scenes.DrawData(width=64, height=64, positions=pos[:600], colors=col[:600], shader=abi.SHADER_INTERPOLATE)
    b = scenes.DrawData(width=64, height=64, positions=pos[600:], colors=col[600:], shader=abi.SHADER_INTERPOLATE,
                        blend_enable=1, depth_compare=abi.COMPARE_LESS_OR_EQUAL,
                        blend=(abi.BLEND_FACTOR_SRC_ALPHA, abi.BLEND_FACTOR_ONE_MINUS_SRC_ALPHA, abi.BLEND_OP_ADD,
                               abi.BLEND_FACTOR_ONE, abi.BLEND_FACTOR_ZERO, abi.BLEND_OP_ADD))
    da = runtime.prepare(a, DEV)
    db = runtime.prepare(b, DEV, targets=da.targets)
    return [n for n, _ in cases] + ["shared_a", "shared_b"], dds + [da, db], draws + [a, b]


def test_batch_matches_oracle_in_order():
    names, dds, draws = mixed_batch()
    print(f"[gpu batch] eager batch of {len(dds)} draws", flush=True)
    runtime.draw_batch(dds)
    torch.cuda.synchronize()
    got = _snap(dds)
    for i, (name, d) in enumerate(zip(names[:-2], draws[:-2])):
        _same(oracle_render(d), got[i], name)
    # The two draws into one target, in order.
    tg = host_targets(draws[-2])
    ra = oracle_render(draws[-2], tg)
    rb = oracle_render(draws[-1], (ra[0], ra[1], ra[2]))
    want = (rb[0], rb[1], rb[2], ra[3] + rb[3])
    _same(want, got[-1], "shared")


def _plan(dds, headroom=10):
    arr = (abi.DrawDesc * len(dds))(*[d.desc for d in dds])
    plan = ctypes.c_void_p()
    runtime.check(runtime.lib().swr_batch_plan_create(arr, len(dds), headroom, ctypes.byref(plan),
                                                      runtime._stream_handle(None)), "plan_create")
    return plan, arr


def test_plan_runs_and_graph_replays_match_eager():
    names, dds, _ = mixed_batch()
    runtime.draw_batch(dds)
    torch.cuda.synchronize()
    want = _snap(dds)
    plan, keep = _plan(dds)
    lib = runtime.lib()
    try:
        _reset(dds)
        runtime.check(lib.swr_batch_plan_run(plan, runtime._stream_handle(None)))
        torch.cuda.synchronize()
        for i, g in enumerate(_snap(dds)):
            _same(want[i], g, names[i])
        # Capture reset + run into one graph; replay with other work between.
        fresh = [runtime.make_targets(dd.draw, DEV) for dd in dds]
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            for dd, f in zip(dds, fresh):
                t = dd.targets
                t.color.copy_(f.color)
                if t.depth is not None:
                    t.depth.copy_(f.depth)
                if t.stencil is not None:
                    t.stencil.copy_(f.stencil)
                t.occlusion.zero_()
            runtime.check(lib.swr_batch_plan_run(plan, runtime._stream_handle(None)))
        for rep in range(4):
            print(f"[gpu batch] graph replay {rep}", flush=True)
            noise = torch.randn(1 << 20, device=DEV)
            (noise * 2).sum().item()
            graph.replay()
            torch.cuda.synchronize()
            for i, g in enumerate(_snap(dds)):
                _same(want[i], g, names[i])
        need = ctypes.c_uint32(0)
        runtime.check(lib.swr_batch_plan_check(plan, ctypes.byref(need)))
        assert need.value == 0
    finally:
        lib.swr_batch_plan_destroy(plan)


def test_every_tuning_renders_the_same_bytes():
    names, dds, _ = mixed_batch()
    lib = runtime.lib()
    base = abi.Tuning()
    runtime.check(lib.swr_get_tuning(ctypes.byref(base)))
    want = None
    try:
        for tw, th, wg in ((16, 16, 4), (16, 16, 0xFFFFFFFF), (16, 8, 6), (8, 8, 12), (8, 16, 2), (8, 8, 0xFFFFFFFF)):
            t = abi.Tuning(tw, th, wg)
            runtime.check(lib.swr_set_tuning(ctypes.byref(t)))
            print(f"[gpu batch] tuning {tw}x{th} wg/cu={wg}", flush=True)
            _reset(dds)
            runtime.draw_batch(dds)
            torch.cuda.synchronize()
            got = _snap(dds)
            if want is None:
                want = got
            else:
                for i, g in enumerate(got):
                    _same(want[i], g, (names[i], tw, th, wg))
    finally:
        runtime.check(lib.swr_set_tuning(ctypes.byref(base)))


def test_autotune_picks_a_candidate_and_keeps_results():
    names, dds, _ = mixed_batch()
    runtime.draw_batch(dds)
    torch.cuda.synchronize()
    want = _snap(dds)
    plan, keep = _plan(dds)
    lib = runtime.lib()
    try:
        best = abi.Tuning()
        runtime.check(lib.swr_batch_plan_autotune(plan, ctypes.byref(best), runtime._stream_handle(None)))
        assert best.tile_width in (8, 16) and best.tile_height in (8, 16)
        _reset(dds)
        runtime.check(lib.swr_batch_plan_run(plan, runtime._stream_handle(None)))
        torch.cuda.synchronize()
        for i, g in enumerate(_snap(dds)):
            _same(want[i], g, names[i])
    finally:
        lib.swr_batch_plan_destroy(plan)
