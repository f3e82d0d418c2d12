#!/usr/bin/env python3
"""Throughput of the fixed-function draw path on one MI355X.

Workload (one step = one frame): a 1920x1080 RGBA8 + D32F target is
cleared (colour 0, depth 1.0) and one draw of 100,000 triangles is
rendered into it -- triangle screen area log-uniform in [1, 1e4] px,
per-vertex w ~ U[0.5, 4], UV ~ U[0, 4], a 1024^2 RGBA8 noise texture with an
11-level box mip chain sampled trilinear with REPEAT, depth LESS, blend
SRC_ALPHA / ONE_MINUS_SRC_ALPHA, texture x colour.  Inputs are resident in
HBM before timing starts.

Reported value: shaded fragments per second, where a fragment is a
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
  cpu_baseline  the project's scalar CPU restatement of the same path
                (oracle/, single thread) on a bounded sample: the first
                40,000 triangles of the same scene at full resolution.
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
    ap.add_argument("--cpu-triangles", type=int, default=40000,
                    help="triangles in the CPU-baseline sample (0 disables it)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 256 MiB device copy (calibrates the HBM byte counters)")
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


def cpu_baseline(draw, n_tri, gpu_count_fn):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_runner import oracle_render  # the CPU oracle: measured here, never used by the GPU path
    sub = dataclasses.replace(draw, positions=draw.positions[:3 * n_tri], colors=draw.colors[:3 * n_tri],
                              texcoords=draw.texcoords[:3 * n_tri])
    frags = gpu_count_fn(sub)
    t0 = time.perf_counter()
    oracle_render(sub)
    dt = time.perf_counter() - t0
    return {"value": frags / dt / 1e9, "unit": "Gfragments/s", "cores": 1, "kind": "port",
            "sample": f"first {n_tri} of the 100,000 triangles at 1920x1080 ({frags} fragments, {dt:.2f} s, "
                      f"scalar C restatement of the reference path, 1 thread)"}


def main():
    args = parse()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(shard whole draws or screen tiles across GPUs: see DESIGN.md)", file=sys.stderr)
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
        color.zero_()
        depth.fill_(1.0)
        occ.zero_()
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
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": traffic_from_profiles(),
                "kernel": "raster_kernel", "kernel_ms": round(raster_ms, 4), "algorithmic_bytes": algo_bytes}

    cpu = None
    if not args.no_cpu and args.cpu_triangles > 0:
        cpu = cpu_baseline(draw, min(args.cpu_triangles, T), covered_fragments)

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
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
