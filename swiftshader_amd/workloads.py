"""Device-resident forms of the benchmark workloads.

The batched workload's textures are generated on the device (512 RGBA8
noise textures of side 256..2048 with full box mip chains would take
minutes to build and upload from the host); this module builds them the
same way for the benchmark and for the tests that compare its draws with
the oracle, so that both see identical bytes.
"""
from typing import List, Tuple

import torch

from . import diff, runtime, scenes


def batched_device(dev, seed: int = 3) -> Tuple[List[scenes.DrawData], List[runtime.DeviceDraw], List[list]]:
    """The 512-draw batched workload with device textures.

    Geometry comes from scenes.batched(seed, host_textures=False); texture
    level 0 of draw k is torch.randint(0, 256) from one device generator
    seeded with `seed`, drawn in draw order, and its mips are the library's
    box chain (swr_generate_mips).  Returns the draws, their prepared
    device draws (targets cleared) and each draw's device texture levels."""
    draws = scenes.batched(seed=seed, host_textures=False)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    dds, levels_all = [], []
    for d in draws:
        ts = d.texture.size
        lvl0 = torch.randint(0, 256, (ts, ts, 4), dtype=torch.uint8, device=dev, generator=gen)
        levels = [lvl0] + diff.mip_levels_like(lvl0)
        diff.generate_mips(levels)
        dds.append(runtime.prepare(d, dev, texture_tensors=levels))
        levels_all.append(levels)
    torch.cuda.synchronize()
    return draws, dds, levels_all
