"""Full-size parity: the benchmark workloads themselves against the oracle.

* the 1920x1080 frame of 100,000 textured triangles that the benchmark
  times (seed 2): colour and depth bytes and the occlusion count equal the
  oracle's, bit for bit;
* the 512-draw batched workload (seed 3, device textures) through one
  autotuned batch plan captured in a HIP graph and replayed twice: 64 of the
  draws, spread over the batch, equal the oracle per draw;
* the 16 differentiable 512^2 views (seed 4) and the 64 uneven 4x MSAA,
  8x anisotropic views (seed 5) through the captured batched fitting step:
  colour, depth and the primitive_id record of every checked view equal the
  oracle's, and the resolved images equal the oracle's resolve.

The oracle runs on several host threads (swo_draw_mt, which renders the
same bytes as swo_draw; tests/test_oracle_mt.py) or several draws at once.
"""
from concurrent.futures import ThreadPoolExecutor
import dataclasses

import numpy as np
import pytest
import torch

from oracle_runner import host_threads, oracle_render, oracle_render_ids, oracle_resolve
from swiftshader_amd import scenes, texfit, workloads
from swiftshader_amd.batch import BatchPlan

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _first_diff(a, b):
    bad = a != b
    if bad.ndim > 3:
        bad = bad.any(-1)
    return int(bad.sum()), (np.argwhere(bad)[0].tolist() if bad.any() else None)


def _assert_same(name, ref, got):
    """ref, got: (colour, depth, occlusion); arrays compared as bytes."""
    rc, rz, ro = ref
    gc, gz, go = got
    n, at = _first_diff(rc.view(np.uint8).reshape(rc.shape[:3] + (-1,)), gc.view(np.uint8).reshape(gc.shape[:3] + (-1,)))
    assert n == 0, f"{name}: {n} pixels of colour differ, first at {at}"
    n, at = _first_diff(rz.view(np.uint32), gz.view(np.uint32))
    assert n == 0, f"{name}: {n} depth values differ, first at {at}"
    assert ro == go, f"{name}: occlusion {go} != oracle {ro}"


def test_1080p_frame_bit_exact():
    """The benchmark's frame: clear + 100k triangles, trilinear RGBA8, blend."""
    from swiftshader_amd import runtime
    draw = scenes.textured_1080p(seed=2)
    print(f"[fullsize] 1080p frame {draw.triangle_count} triangles", flush=True)
    dd = runtime.prepare(draw, DEV)
    runtime.clear_targets(draw, dd.targets)
    runtime.draw(dd)
    torch.cuda.synchronize()
    got = (dd.targets.color.cpu().numpy(), dd.targets.depth.cpu().numpy(), int(dd.targets.occlusion.item()))
    rc, rz, _, ro = oracle_render(draw, threads=host_threads())
    print(f"[fullsize] 1080p oracle done, occlusion {ro}", flush=True)
    _assert_same("1080p", (rc, rz, ro), got)
    assert ro > 9_000_000


def test_batched_512_draws_plan_graph():
    """The exact benchmark batch (autotuned plan, clears + run in one graph,
    replayed twice); 64 draws spread over the batch against the oracle."""
    from swiftshader_amd import runtime
    draws, dds, levels = workloads.batched_device(DEV, seed=3)
    print(f"[fullsize] batch of {len(dds)} draws, {sum(d.triangle_count for d in draws)} triangles", flush=True)
    plan = BatchPlan([dd.desc for dd in dds])
    plan.autotune()

    def clear_all():
        for dd in dds:
            runtime.clear_targets(dd.draw, dd.targets, reset_occlusion=True)

    clear_all()
    plan.run()
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        clear_all()
        plan.run()
    for _ in range(2):
        graph.replay()
        torch.cuda.synchronize()
    assert plan.pairs_overflow() == 0
    picks = list(range(0, len(draws), 8))
    got = {k: (dds[k].targets.color.cpu().numpy(), dds[k].targets.depth.cpu().numpy(),
               int(dds[k].targets.occlusion.item())) for k in picks}
    host = {k: dataclasses.replace(draws[k], texture=dataclasses.replace(
        draws[k].texture, levels=[l.cpu().numpy() for l in levels[k]])) for k in picks}
    plan.close()

    def ref(k):
        rc, rz, _, ro = oracle_render(host[k])
        return k, (rc, rz, ro)

    with ThreadPoolExecutor(max_workers=host_threads()) as ex:
        for k, r in ex.map(ref, picks):
            _assert_same(f"draw {k} ({draws[k].width}x{draws[k].height}, {draws[k].triangle_count} tris)", r, got[k])
    print(f"[fullsize] {len(picks)} draws bit-exact", flush=True)


def _fit_views_against_oracle(draws, views_to_check, label):
    fit = texfit.MultiViewFit(draws, torch.from_numpy(draws[0].texture.levels[0]), device=DEV)
    fit.capture(update=False)
    fit.step_graph()
    fit.step_graph()
    torch.cuda.synchronize()
    assert fit.plan.pairs_overflow() == 0
    levels = [l.cpu().numpy() for l in fit.levels]
    # The device chain is the box chain of level 0 (tests/test_gpu_grad.py);
    # the oracle samples exactly the device's levels.
    tex = dataclasses.replace(draws[0].texture, levels=levels)
    got = {}
    for k in views_to_check:
        vb = fit.views[k]
        got[k] = (vb.color.cpu().numpy(), vb.depth.cpu().numpy(), vb.ids.cpu().numpy().view(np.uint32),
                  vb.resolved.cpu().numpy() if vb.draw.samples > 1 else None)
    del fit
    torch.cuda.empty_cache()

    def ref(k):
        d = dataclasses.replace(draws[k], texture=tex)
        return k, oracle_render_ids(d)

    with ThreadPoolExecutor(max_workers=host_threads()) as ex:
        for k, (rc, rz, rids, _) in ex.map(ref, views_to_check):
            gc, gz, gids, gres = got[k]
            name = f"{label} view {k} ({draws[k].width}x{draws[k].height} s={draws[k].samples})"
            n, at = _first_diff(rc.view(np.uint32), gc.view(np.uint32))
            assert n == 0, f"{name}: {n} pixels of colour differ, first at {at}"
            n, at = _first_diff(rz.view(np.uint32), gz.view(np.uint32))
            assert n == 0, f"{name}: {n} depth values differ, first at {at}"
            n, at = _first_diff(rids, gids)
            assert n == 0, f"{name}: {n} primitive ids differ, first at {at}"
            if gres is not None:
                want = oracle_resolve(rc)
                assert np.array_equal(want.view(np.uint32), gres.view(np.uint32)), f"{name}: resolve differs"
    print(f"[fullsize] {label}: {len(views_to_check)} views bit-exact", flush=True)


def test_diff_views_16_batched_forward():
    draws = scenes.diff_views(seed=4)
    print(f"[fullsize] 16 differentiable views", flush=True)
    _fit_views_against_oracle(draws, list(range(len(draws))), "diff_views")


def test_msaa_aniso_64_batched_forward():
    draws = scenes.msaa_aniso(seed=5)
    print(f"[fullsize] 64 MSAA aniso views", flush=True)
    order = sorted(range(len(draws)), key=lambda k: draws[k].width * draws[k].height)
    picks = sorted(set(list(range(0, len(draws), 6)) + order[-2:]))
    _fit_views_against_oracle(draws, picks, "msaa_aniso")
