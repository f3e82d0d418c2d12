"""GPU parity: the HIP draw path against the CPU oracle, bit for bit.

Every case renders the same seeded scene through libswr.so (C ABI) on the
GPU and through the oracle on the CPU and requires identical colour,
depth, stencil bytes and occlusion count (looser tolerances -- depth
2^-22 relative, UNORM8 1 LSB -- are not needed: the kernels reproduce the
reference's arithmetic exactly).
"""
import sys

import numpy as np
import pytest

from cases import parity_cases
from oracle_runner import oracle_render, oracle_resolve

pytestmark = pytest.mark.gpu

CASES = parity_cases()


def gpu_render(draw):
    import torch
    from swiftshader_amd import runtime
    dd = runtime.prepare(draw, "cuda:0")
    runtime.draw(dd)
    torch.cuda.synchronize()
    t = dd.targets
    color = t.color.cpu().numpy()
    depth = t.depth.cpu().numpy() if t.depth is not None else None
    stencil = t.stencil.cpu().numpy() if t.stencil is not None else None
    return color, depth, stencil, int(t.occlusion.cpu().item())


def _compare(name, ref, got):
    rc, rz, rs, ro = ref
    gc, gz, gs, go = got
    if rc.dtype == np.float32:
        bad = (rc.view(np.uint32) != gc.view(np.uint32))
    else:
        bad = rc != gc
    assert not bad.any(), f"{name}: {int(bad.sum())} colour components differ; first at {np.argwhere(bad)[0]}"
    if rz is not None:
        badz = rz.view(np.uint8).reshape(rz.shape + (-1,)) != gz.view(np.uint8).reshape(gz.shape + (-1,))
        badz = badz.any(-1)
        assert not badz.any(), f"{name}: {int(badz.sum())} depth values differ; first at {np.argwhere(badz)[0]}"
    if rs is not None:
        assert np.array_equal(rs, gs), f"{name}: stencil differs"
    assert ro == go, f"{name}: occlusion {go} != oracle {ro}"


@pytest.mark.parametrize("name,draw", CASES, ids=[c[0] for c in CASES])
def test_draw_matches_oracle(name, draw):
    print(f"[gpu parity] {name} {draw.width}x{draw.height} samples={draw.samples} tris={draw.triangle_count}",
          flush=True)
    sys.stdout.flush()
    ref = oracle_render(draw)
    got = gpu_render(draw)
    _compare(name, ref, got)


def test_resolve_matches_oracle():
    import torch
    from swiftshader_amd import runtime
    rng = np.random.default_rng(5)
    for dt in (np.uint8, np.float32):
        src = (rng.integers(0, 256, (4, 33, 47, 4)).astype(np.uint8) if dt == np.uint8
               else rng.normal(0.5, 0.3, (4, 33, 47, 4)).astype(np.float32))
        print(f"[gpu parity] resolve {src.dtype}", flush=True)
        ref = oracle_resolve(src)
        s = torch.from_numpy(src).cuda()
        d = torch.zeros((1, 33, 47, 4), dtype=s.dtype, device="cuda")
        runtime.resolve(s, d)
        torch.cuda.synchronize()
        got = d.cpu().numpy()
        assert np.array_equal(ref.view(np.uint8), got.view(np.uint8))


def test_draw_beyond_the_known_pair_capacity():
    """After the first single draw, swr_draw stops waiting on the host: it
    sizes the (tile, primitive) pair buffers from the counts of earlier draws
    and renders the pairs beyond that capacity through the unbinned pass.  A
    large draw right after a small one (far beyond the small one's capacity),
    then the same draw again (capacity grown from its count), both match the
    oracle bit for bit."""
    import torch
    from swiftshader_amd import runtime
    # The smallest case leaves a capacity of at most a few thousand pairs
    # (its count plus an eighth plus 1024); the case with the most triangles
    # needs at least one pair per triangle, well beyond it.
    by_size = sorted(CASES, key=lambda c: (c[1].triangle_count, c[1].width * c[1].height))
    small, big = by_size[0], by_size[-1]
    assert small[1].triangle_count <= 8 and big[1].triangle_count >= 3000
    runtime.check(runtime.lib().swr_release(), "swr_release")  # forget the learnt capacity
    print(f"[gpu parity] capacity from {small[0]}, then {big[0]} twice", flush=True)
    gpu_render(small[1])
    ref = oracle_render(big[1])
    _compare(big[0] + " (beyond capacity)", ref, gpu_render(big[1]))
    torch.cuda.synchronize()
    _compare(big[0] + " (capacity grown)", ref, gpu_render(big[1]))


def _oracle_threads():
    import os
    return max(1, min(16, os.cpu_count() or 1))


def test_draw_above_the_one_launch_scan_limit():
    """More than 1024 x 1024 primitive slots: the tile-count scan takes its
    three-launch form (block sums scanned by one workgroup) and the emission
    index its own kernel, both in the first draw (which reads the pair count
    back) and in the second (whose pair capacity is known beforehand).  Both
    match the oracle bit for bit."""
    import torch
    from swiftshader_amd import runtime, scenes
    draw = scenes.plain_gouraud(seed=11, triangles=1_100_000, size=64)
    runtime.check(runtime.lib().swr_release(), "swr_release")  # the first draw waits for its count
    print(f"[gpu parity] 1.1M triangles at 64x64, twice", flush=True)
    ref = oracle_render(draw, threads=_oracle_threads())
    _compare("1.1M triangles (count read back)", ref, gpu_render(draw))
    torch.cuda.synchronize()
    _compare("1.1M triangles (capacity known)", ref, gpu_render(draw))


def test_draw_above_the_fused_radix_limit():
    """More than 4096 sort blocks of (tile, primitive) pairs (8x8 tiles, 250k
    triangles spanning about a quarter of a 128x128 target: about 20 M pairs):
    the radix passes scan their histograms in three launches instead of
    adding the block totals in the scatter.  The first draw sizes its pair
    buffers from the count it reads back, the second from the capacity learnt
    from it; both sort every pair and match the oracle bit for bit."""
    import ctypes
    import torch
    from swiftshader_amd import abi, runtime, scenes
    lib = runtime.lib()
    base = abi.Tuning()
    runtime.check(lib.swr_get_tuning(ctypes.byref(base)))
    draw = scenes.plain_gouraud(seed=12, triangles=250_000, size=128)
    print(f"[gpu parity] 250k triangles at 128x128, 8x8 tiles", flush=True)
    ref = oracle_render(draw, threads=_oracle_threads())
    try:
        runtime.check(lib.swr_set_tuning(ctypes.byref(abi.Tuning(8, 8, 4))))
        runtime.check(lib.swr_release(), "swr_release")  # no capacity learnt from smaller draws
        _compare("250k triangles, 8x8 tiles (count read back)", ref, gpu_render(draw))
        torch.cuda.synchronize()
        _compare("250k triangles, 8x8 tiles (capacity known)", ref, gpu_render(draw))
        torch.cuda.synchronize()
        runtime.profile_read()  # discard what earlier draws left
        runtime.profile_enable(True)
        _compare("250k triangles, 8x8 tiles (profiled)", ref, gpu_render(draw))
        prof = runtime.profile_read()
        assert prof["draws"] == 1 and prof["pairs"] > 4096 * 4096, prof
    finally:
        runtime.profile_enable(False)
        runtime.check(lib.swr_set_tuning(ctypes.byref(base)))
