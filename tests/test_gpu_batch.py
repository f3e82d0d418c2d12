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


def _hazard_draws():
    """A renders into image IA; B samples IA as its texture; C and D render
    blended into two overlapping 48-column windows of one 64x64 image (base
    pointers 16 pixels apart, one row pitch)."""
    from swiftshader_amd.scenes import DrawData, TextureData
    rng = np.random.default_rng(21)
    pos, uv, col = scenes.screen_triangles(rng, 300, 64, 64, area_hi=500)
    a = DrawData(width=64, height=64, positions=pos[:450], colors=col[:450], shader=abi.SHADER_INTERPOLATE,
                 clear_color=(0.1, 0.2, 0.3, 1.0))
    b = DrawData(width=64, height=64, positions=pos[450:], texcoords=uv[450:] * 0.5, shader=abi.SHADER_SAMPLE,
                 texture=TextureData(format=abi.FORMAT_R8G8B8A8_UNORM, levels=[np.zeros((64, 64, 4), np.uint8)],
                                     mipmap_mode=abi.MIPMAP_MODE_NEAREST))
    pos2, _, col2 = scenes.screen_triangles(rng, 200, 48, 64, area_hi=400)
    blend = (abi.BLEND_FACTOR_SRC_ALPHA, abi.BLEND_FACTOR_ONE_MINUS_SRC_ALPHA, abi.BLEND_OP_ADD,
             abi.BLEND_FACTOR_ONE, abi.BLEND_FACTOR_ZERO, abi.BLEND_OP_ADD)
    c = DrawData(width=48, height=64, positions=pos2[:300], colors=col2[:300], shader=abi.SHADER_INTERPOLATE,
                 has_depth=False, depth_test=0, blend_enable=1, blend=blend)
    d = DrawData(width=48, height=64, positions=pos2[300:], colors=col2[300:], shader=abi.SHADER_INTERPOLATE,
                 has_depth=False, depth_test=0, blend_enable=1, blend=blend)
    return a, b, c, d


def _hazard_device(a, b, c, d):
    da = runtime.prepare(a, DEV)
    db = runtime.prepare(b, DEV, texture_tensors=[da.targets.color[0]])
    image = torch.zeros((1, 64, 64, 4), dtype=torch.uint8, device=DEV)
    dc = runtime.prepare(c, DEV)
    dd = runtime.prepare(d, DEV)
    for dev_draw, x0 in ((dc, 0), (dd, 16)):
        dev_draw.desc.color = abi.Attachment(image.data_ptr() + 4 * x0, abi.FORMAT_R8G8B8A8_UNORM, 64 * 4, 64 * 64 * 4)
    return [da, db, dc, dd], image


def test_batch_orders_draws_that_read_or_overlap_earlier_writes():
    """A batch equals its draws issued one by one on one stream (and the
    oracle) when a draw samples an earlier draw's colour attachment and when
    two attachments overlap without sharing a base pointer."""
    a, b, c, d = _hazard_draws()
    seq, img_seq = _hazard_device(a, b, c, d)
    for x in seq:
        runtime.draw(x)
    bat, img_bat = _hazard_device(a, b, c, d)
    print("[gpu batch] hazard batch", flush=True)
    runtime.draw_batch(bat)
    torch.cuda.synchronize()
    assert np.array_equal(seq[0].targets.color.cpu().numpy(), bat[0].targets.color.cpu().numpy())
    assert np.array_equal(seq[1].targets.color.cpu().numpy(), bat[1].targets.color.cpu().numpy())
    assert np.array_equal(img_seq.cpu().numpy(), img_bat.cpu().numpy())
    # The oracle in API order.
    ra = oracle_render(a)
    import dataclasses
    b_host = dataclasses.replace(b, texture=dataclasses.replace(b.texture, levels=[ra[0][0].copy()]))
    rb = oracle_render(b_host)
    assert np.array_equal(rb[0], bat[1].targets.color.cpu().numpy())
    img = np.zeros((1, 64, 64, 4), np.uint8)
    for x0, dr in ((0, c), (16, d)):
        win = np.ascontiguousarray(img[:, :, x0:x0 + 48])
        r = oracle_render(dr, (win, None, None))
        img[:, :, x0:x0 + 48] = r[0]
    assert np.array_equal(img, img_bat.cpu().numpy())
    # The same batch through a plan.
    plan, keep = _plan(bat)
    lib = runtime.lib()
    try:
        fresh, img2 = _hazard_device(a, b, c, d)
        for x, f in zip(bat, fresh):
            x.targets.color.copy_(f.targets.color)
            if x.targets.depth is not None:
                x.targets.depth.copy_(f.targets.depth)
        img_bat.zero_()
        runtime.check(lib.swr_batch_plan_run(plan, runtime._stream_handle(None)))
        torch.cuda.synchronize()
        assert np.array_equal(rb[0], bat[1].targets.color.cpu().numpy())
        assert np.array_equal(img, img_bat.cpu().numpy())
    finally:
        lib.swr_batch_plan_destroy(plan)


def test_plan_whose_geometry_outgrows_it_still_renders_every_primitive():
    """A plan built with no headroom, whose vertices then move so that the
    draws need several times the planned (tile, primitive) pairs: the run
    (eager and replayed from a graph) still equals the oracle bit for bit,
    and swr_batch_plan_check reports the pairs the run needed."""
    from swiftshader_amd.scenes import DrawData
    rng = np.random.default_rng(31)
    small, big = [], []
    for k, (w, h, samples) in enumerate(((96, 80, 1), (128, 96, 4), (64, 64, 1))):
        ps, _, cs = scenes.screen_triangles(rng, 600, w, h, area_lo=2.0, area_hi=20.0)
        pb, _, cb = scenes.screen_triangles(rng, 600, w, h, area_lo=200.0, area_hi=2000.0)
        kw = dict(width=w, height=h, samples=samples, shader=abi.SHADER_INTERPOLATE, blend_enable=1,
                  blend=(abi.BLEND_FACTOR_SRC_ALPHA, abi.BLEND_FACTOR_ONE_MINUS_SRC_ALPHA, abi.BLEND_OP_ADD,
                         abi.BLEND_FACTOR_ONE, abi.BLEND_FACTOR_ZERO, abi.BLEND_OP_ADD))
        small.append(DrawData(positions=ps, colors=cs, **kw))
        big.append(DrawData(positions=pb, colors=cb, **kw))
    dds = [runtime.prepare(d, DEV) for d in small]
    plan, keep = _plan(dds, headroom=0)
    lib = runtime.lib()
    try:
        # Same counts, new vertex contents: the plan reads them at run time.
        for dd, b in zip(dds, big):
            dd.tensors[0].copy_(torch.from_numpy(b.positions))
            dd.tensors[1].copy_(torch.from_numpy(b.colors))
        want = [oracle_render(b) for b in big]
        _reset(dds)
        runtime.check(lib.swr_batch_plan_run(plan, runtime._stream_handle(None)))
        torch.cuda.synchronize()
        for i, g in enumerate(_snap(dds)):
            _same(want[i], g, f"grown draw {i}")
        need = ctypes.c_uint32(0)
        runtime.check(lib.swr_batch_plan_check(plan, ctypes.byref(need)))
        assert need.value > 0
        fresh = [runtime.make_targets(dd.draw, DEV) for dd in dds]
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            for dd, f in zip(dds, fresh):
                dd.targets.color.copy_(f.color)
                dd.targets.depth.copy_(f.depth)
                dd.targets.occlusion.zero_()
            runtime.check(lib.swr_batch_plan_run(plan, runtime._stream_handle(None)))
        for rep in range(2):
            print(f"[gpu batch] grown plan graph replay {rep}", flush=True)
            graph.replay()
            torch.cuda.synchronize()
            for i, g in enumerate(_snap(dds)):
                _same(want[i], g, f"grown draw {i} replay {rep}")
    finally:
        lib.swr_batch_plan_destroy(plan)
