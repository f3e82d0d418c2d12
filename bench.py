#!/usr/bin/env python3
"""Throughput of the fixed-function draw path on one MI355X.

Workload (one step = one frame): a 1920x1080 RGBA8 + D32F target is
cleared (colour 0, depth 1.0) and one draw of 100,000 triangles is
rendered into it -- triangle screen area log-uniform in [1, 1e4] px,
per-vertex w ~ U[0.5, 4], UV ~ U[0, 4], a 1024^2 RGBA8 noise texture with an
11-level box mip chain sampled trilinear with REPEAT, depth LESS, blend
SRC_ALPHA / ONE_MINUS_SRC_ALPHA, texture x colour.  Inputs are resident in
HBM before timing starts.

Reported value (headline): shaded fragments per second, where a fragment is a
(pixel, primitive) pair with non-empty coverage after the scissor -- the
count is taken with an occlusion query of a depth-ALWAYS counting draw of
the same scene, outside the timed region.

Also reported:
  roofline      the tile rasteriser's algorithmic bytes per launch
                (per fragment: 32 B trilinear texel footprint + 4 B depth
                read; per depth-passing fragment: 4 B depth write + 4 B
                colour write + 4 B colour read for the blend; 3 x 40 B of
                vertex data per triangle; 16 B per pixel of tile load and
                store) divided by its HIP-event-timed duration, against
                the 8 TB/s HBM peak.
  cpu_baseline  the project's CPU restatement of the same path (oracle/),
                on every host core this process may use (at most 16): row
                pairs split over threads as the reference splits them over
                its rasteriser clusters, primitives set up in batches by
                all threads -- the whole 100,000-triangle frame.
  batched       512 uneven draws: frames (draws) per second of one
                autotuned batch plan replayed from a HIP graph, against one
                swr_draw call per draw.
  fwd_bwd       16 differentiable 512^2 views: ms per forward + backward
                step, graph-captured batch against per-draw launches.
  texture_fit   64-view 4x MSAA, 8x anisotropic texture fitting with Adam:
                ms per captured step against per-draw launches, and the loss
                before and after.
"""
import argparse
import dataclasses
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBPS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-triangles", type=int, default=100000,
                    help="triangles of the frame in the CPU-baseline sample (0 disables it)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU-baseline threads (0: the usable cores, <= 16)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 256 MiB device copy (calibrates the HBM byte counters)")
    ap.add_argument("--legs", default="draw,batched,fwd_bwd,fit",
                    help="comma-separated subset of: draw (1080p frame, the headline), batched (512 uneven draws), "
                         "fwd_bwd (16 differentiable views), fit (64-view 4x MSAA texture fitting)")
    ap.add_argument("--fit-steps", type=int, default=200, help="Adam steps of the fitting leg")
    return ap.parse_args()


def copy_probe():
    """A plain device-to-device copy of known size, for scaling FETCH_SIZE / WRITE_SIZE."""
    import torch
    n = 256 << 20
    src = torch.empty(n, dtype=torch.uint8, device="cuda:0").fill_(1)
    dst = torch.empty_like(src)
    for _ in range(4):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n, "copy_probe_bytes_written": n, "launches": 4}), flush=True)


def traffic_from_profiles():
    """HBM bytes per raster launch from the committed rocprofv3 PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", "raster_textured_1080p_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def cpu_baseline(draw, n_tri, gpu_count_fn, threads):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    # The CPU oracle: measured here, never used by the GPU path.
    from oracle_runner import host_threads, oracle_render
    threads = threads or host_threads()
    sub = dataclasses.replace(draw, positions=draw.positions[:3 * n_tri], colors=draw.colors[:3 * n_tri],
                              texcoords=draw.texcoords[:3 * n_tri])
    frags = gpu_count_fn(sub)
    oracle_render(dataclasses.replace(sub, positions=sub.positions[:3000], colors=sub.colors[:3000],
                                      texcoords=sub.texcoords[:3000]), threads=threads)  # warm-up
    t0 = time.perf_counter()
    oracle_render(sub, threads=threads)
    dt = time.perf_counter() - t0
    what = "the whole frame" if n_tri == draw.triangle_count else f"the first {n_tri} triangles"
    return {"value": frags / dt / 1e9, "unit": "Gfragments/s", "cores": threads, "kind": "port",
            "sample": f"{what} ({n_tri} of {draw.triangle_count} triangles) at 1920x1080, {frags} fragments in "
                      f"{dt:.2f} s: C restatement of the reference path (oracle/), row pairs split over {threads} "
                      f"threads as the reference's clusters split them"}


def _ms(fn, reps):
    """Average milliseconds of fn() over reps, HIP events on the current stream."""
    import torch
    st = torch.cuda.current_stream()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record(st)
    for _ in range(reps):
        fn()
    b.record(st)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def leg_batched(dev, reps):
    """512 independent draws of uneven size (targets 64..1024 per side, 2k..50k
    triangles, RGBA8 noise textures 256..2048 with full mips, trilinear,
    CLAMP_TO_EDGE or REPEAT): one swr_clear_batch of every target and one
    autotuned batch plan run, captured in one HIP graph, against one swr_clear
    per target and one swr_draw per draw (no graph)."""
    import torch
    from swiftshader_amd import runtime, workloads
    from swiftshader_amd.batch import BatchPlan
    draws, dds, _ = workloads.batched_device(dev, seed=3)
    tris = sum(d.triangle_count for d in draws)

    def clear_all():
        runtime.clear_targets_batch([(dd.draw, dd.targets) for dd in dds])

    def plain():
        for dd in dds:
            runtime.clear_targets(dd.draw, dd.targets, reset_occlusion=False)
            runtime.draw(dd)

    plain()
    plain_ms = _ms(plain, 2)
    plan = BatchPlan([dd.desc for dd in dds])
    best = plan.autotune()
    clear_all()
    plan.run()
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        clear_all()
        plan.run()
    graph.replay()
    batched_ms = _ms(graph.replay, reps)
    over = plan.pairs_overflow()
    plan.close()
    return {"workload": "512 draws, targets 64..1024 px per side, 2k..50k triangles each "
                        f"({tris} total), RGBA8 textures 256..2048 with mips, trilinear, SRC_ALPHA blend",
            "draws": len(dds), "batch_ms": round(batched_ms, 3), "frames_per_s": round(1e3 * len(dds) / batched_ms, 1),
            "plain_batch_ms": round(plain_ms, 3), "plain_frames_per_s": round(1e3 * len(dds) / plain_ms, 1),
            "speedup_vs_per_draw": round(plain_ms / batched_ms, 2),
            "tuning": {"tile": f"{best.tile_width}x{best.tile_height}",
                       "raster_workgroups_per_cu": ("one per tile" if best.raster_workgroups_per_cu == 0xFFFFFFFF
                                                    else int(best.raster_workgroups_per_cu))},
            "pair_overflow": over}


def _fit_targets(draws, tex_levels0, dev):
    """Target images: the views rendered with a second texture."""
    import torch
    from swiftshader_amd import texfit
    ref = texfit.MultiViewFit(draws, torch.from_numpy(tex_levels0), device=dev)
    ref.forward_plain()
    imgs = []
    for vb in ref.views:
        if vb.draw.samples > 1:
            from swiftshader_amd import runtime
            runtime.resolve(vb.color, vb.resolved)
            imgs.append(vb.resolved[0].clone())
        else:
            imgs.append(vb.color[0].clone())
    torch.cuda.synchronize()
    del ref
    return imgs


def leg_fwd_bwd(dev, reps):
    """16 views 512^2 RGBA32F, 20k triangles each, 1024^2 float texture
    (trilinear): forward + backward (texel, colour, texcoord and position
    gradients) of an L2 loss, per-draw launches vs one graph-captured step."""
    import numpy as np
    import torch
    from swiftshader_amd import scenes, texfit
    draws = scenes.diff_views(seed=4)
    rng = np.random.default_rng(40)
    targets = _fit_targets(draws, scenes.float_texture(rng, 1024)[0], dev)
    fit = texfit.MultiViewFit(draws, torch.from_numpy(draws[0].texture.levels[0]), targets,
                              geometry_grads=True, device=dev)
    fit.step_plain(update=False)
    plain_ms = _ms(lambda: fit.step_plain(update=False), 3)
    fit.capture(update=False)
    fit.step_graph()
    graph_ms = _ms(fit.step_graph, reps)
    return {"workload": "16 views 512x512 RGBA32F + D32F, 20k triangles each, 1024^2 RGBA32F texture, trilinear; "
                        "L2 loss; gradients of texels, colours, texcoords, positions",
            "plain_ms_per_step": round(plain_ms, 3), "graph_ms_per_step": round(graph_ms, 3),
            "speedup": round(plain_ms / graph_ms, 2), "loss": fit.current_loss(),
            "pair_overflow": fit.plan.pairs_overflow(),
            "gradient_tolerance": "<= 1e-4 relative to the FP64 oracle (tests/test_gpu_grad.py)"}


def leg_fit(dev, steps):
    """64 views of side 256..1024, 4x MSAA resolved by averaging, 10k..40k
    triangles, 2048^2 RGBA32F texture with 8x anisotropic trilinear
    filtering: Adam (lr 1e-2) on the texture against views rendered from a
    second texture; each step captured in one HIP graph."""
    import numpy as np
    import torch
    from swiftshader_amd import scenes, texfit
    draws = scenes.msaa_aniso(seed=5)
    rng = np.random.default_rng(50)
    targets = _fit_targets(draws, scenes.float_texture(rng, 2048)[0], dev)
    fit = texfit.MultiViewFit(draws, torch.from_numpy(draws[0].texture.levels[0]), targets, lr=1e-2, device=dev)
    fit.step_plain(update=False)
    plain_ms = _ms(lambda: fit.step_plain(update=False), 2)
    fit.capture(update=True)
    fit.step_graph()
    torch.cuda.synchronize()
    loss0 = fit.current_loss()
    graph_ms = _ms(fit.step_graph, max(steps - 1, 1))
    loss1 = fit.current_loss()
    return {"workload": "64 views 256..1024 px per side, 4x MSAA + resolve, 10k..40k triangles each, "
                        "2048^2 RGBA32F texture, 8x anisotropic trilinear, Adam lr 1e-2",
            "steps": steps, "plain_ms_per_step": round(plain_ms, 3), "graph_ms_per_step": round(graph_ms, 3),
            "speedup": round(plain_ms / graph_ms, 2), "loss_first_step": loss0, "loss_last_step": loss1,
            "pair_overflow": fit.plan.pairs_overflow()}


def main():
    args = parse()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(whole draws or screen tiles would shard across GPUs)", file=sys.stderr)
        sys.exit(2)

    import torch
    from swiftshader_amd import abi, runtime, scenes

    dev = "cuda:0"
    draw = scenes.textured_1080p(seed=2)
    T = draw.triangle_count
    W, H = draw.width, draw.height

    def covered_fragments(d):
        cnt = dataclasses.replace(d, depth_compare=abi.COMPARE_ALWAYS, depth_write=0, color_write_mask=0)
        dd = runtime.prepare(cnt, dev)
        runtime.draw(dd)
        torch.cuda.synchronize()
        return int(dd.targets.occlusion.item())

    F = covered_fragments(draw)
    dd = runtime.prepare(draw, dev)
    color, depth, occ = dd.targets.color, dd.targets.depth, dd.targets.occlusion

    def step():
        runtime.clear_targets(draw, dd.targets)  # render-pass clear: swr_clear of colour and depth
        runtime.draw(dd)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    P = int(occ.item())  # depth-passing fragments of one frame
    stream = torch.cuda.current_stream()

    def timed_steps():
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        start.record(stream)
        for _ in range(args.steps):
            step()
        end.record(stream)
        torch.cuda.synchronize()
        return start.elapsed_time(end) / args.steps

    # Throughput: the K timed frames, no instrumentation.
    ms_per_step = timed_steps()
    # Stage breakdown and the raster kernel's own duration: the same K frames
    # again with the library's HIP events recorded between the stages (on the
    # stream the kernels run on).
    runtime.profile_read()
    runtime.profile_enable(True)
    ms_profiled = timed_steps()
    runtime.profile_enable(False)
    prof = runtime.profile_read()
    raster_ms = prof["raster_ms"] / max(prof["draws"], 1)

    algo_bytes = F * (32 + 4) + P * (4 + 4 + 4) + 3 * T * 40 + W * H * 16
    achieved = algo_bytes / (raster_ms * 1e-3) / 1e9
    shaded = prof["shaded"] // max(prof["draws"], 1)
    # The formula above prices a 32-byte trilinear footprint for every
    # covered fragment, as the reference fetches it (it shades before the
    # depth test); this kernel tests depth first and samples only for the
    # fragments it shades.  Two further readings of the same launch:
    shaded_bytes = shaded * 32 + F * 4 + P * (4 + 4 + 4) + 3 * T * 40 + W * H * 16
    traffic = traffic_from_profiles()
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": traffic,
                "kernel": "raster_kernel", "kernel_ms": round(raster_ms, 4), "algorithmic_bytes": algo_bytes,
                "frac_footprint_of_shaded_fragments_only": round(shaded_bytes / (raster_ms * 1e-3) / 1e9 /
                                                                 PEAK_HBM_GBPS, 4),
                "frac_measured_traffic": (round(traffic / (raster_ms * 1e-3) / 1e9 / PEAK_HBM_GBPS, 4)
                                          if traffic else None),
                "note": "frac prices the reference's work (a texel footprint per covered fragment); the kernel "
                        "samples only depth-passing fragments and keeps tiles on chip, so it is bound by "
                        "instruction issue, not HBM"}

    cpu = None
    if not args.no_cpu and args.cpu_triangles > 0:
        cpu = cpu_baseline(draw, min(args.cpu_triangles, T), covered_fragments, args.cpu_threads)

    out = {
        "metric": "shaded Gfragments/s",
        "value": round(F / (ms_per_step * 1e-3) / 1e9, 3),
        "unit": "Gfragments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+u8",
        "data": "synthetic (seeded)",
        "config": {"workload": "1920x1080 RGBA8+D32F clear + one draw of 100k textured triangles "
                               "(trilinear RGBA8 1024^2, REPEAT, depth LESS, SRC_ALPHA blend)",
                   "fragments_per_step": F, "depth_passing_per_step": P, "triangles": T,
                   "ms_per_step_with_stage_events": round(ms_profiled, 4),
                   "stage_ms_per_step": {k: round(prof[k] / max(prof["draws"], 1), 4)
                                         for k in ("setup_ms", "bin_ms", "raster_ms", "total_ms")},
                   "pairs_per_step": prof["pairs"] // max(prof["draws"], 1),
                   "block_tests_per_step": prof["block_tests"] // max(prof["draws"], 1),
                   "block_covers_per_step": prof["block_covers"] // max(prof["draws"], 1),
                   "shaded_per_step": prof["shaded"] // max(prof["draws"], 1)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    legs = [x.strip() for x in args.legs.split(",") if x.strip()]
    # The other two parts of the metric (batched frames/s; fwd+bwd ms/step
    # against per-draw launches); each leg prints a progress line to stderr.
    for name, fn in (("batched", lambda: leg_batched(dev, max(args.steps // 4, 3))),
                     ("fwd_bwd", lambda: leg_fwd_bwd(dev, max(args.steps // 4, 3))),
                     ("texture_fit", lambda: leg_fit(dev, args.fit_steps))):
        key = {"batched": "batched", "fwd_bwd": "fwd_bwd", "texture_fit": "fit"}[name]
        if key in legs:
            print(f"[bench] leg {name}", file=sys.stderr, flush=True)
            try:
                out[name] = fn()
            except Exception as e:  # a failing leg is reported, the headline line still prints
                out[name] = {"error": f"{type(e).__name__}: {e}"}
            torch.cuda.synchronize()
            torch.cuda.empty_cache()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
