"""swr_decode_bc on the GPU against the oracle, bit for bit: BC1-BC5 in
every alpha / sign mode, ragged and large images, pixel strides that leave
bytes untouched; and the refused cases (BC6H / BC7, misaligned BC1-BC3)."""
import ctypes

import numpy as np
import pytest
import torch

from oracle_runner import bc_block_bytes, oracle_decode_bc
from swiftshader_amd import abi, runtime

pytestmark = pytest.mark.gpu

SHAPES = [(16, 12), (13, 7), (1, 1), (257, 130), (1024, 768)]


def _gpu_decode(src, w, h, bpp, n, flag, pitch, fill):
    s = torch.from_numpy(src).to("cuda:0")
    d = torch.full((pitch * h,), fill, dtype=torch.uint8, device="cuda:0")
    desc = abi.BCDesc(s.data_ptr(), d.data_ptr(), w, h, pitch, bpp, n, flag)
    rc = runtime.lib().swr_decode_bc(ctypes.byref(desc), None)
    torch.cuda.synchronize()
    return rc, d.cpu().numpy()


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("flag", [0, 1])
def test_decode_matches_the_oracle(n, flag):
    rng = np.random.default_rng(n * 2 + flag)
    for bpp in ((4,) if n <= 3 else ((1, 3) if n == 4 else (2, 4))):
        for w, h in SHAPES:
            print(f"[gpu bc] BC{n} flag {flag} bpp {bpp} {w}x{h}", flush=True)
            nb = ((w + 3) // 4) * ((h + 3) // 4)
            src = rng.integers(0, 256, nb * bc_block_bytes(n), dtype=np.uint8)
            pitch = w * bpp + (4 if n <= 3 else 3)
            want = np.full(pitch * h, 0x5A, np.uint8)
            assert oracle_decode_bc(src, want, w, h, pitch, bpp, n, flag) == 1
            rc, got = _gpu_decode(src, w, h, bpp, n, flag, pitch, 0x5A)
            assert rc == 0 and np.array_equal(got, want), (w, h, bpp)


def test_refused_cases():
    src = np.zeros(64, np.uint8)
    rc, _ = _gpu_decode(src, 4, 4, 4, 7, 0, 16, 0)
    assert rc == abi.ERROR_UNSUPPORTED
    rc, _ = _gpu_decode(src, 4, 4, 4, 6, 0, 16, 0)
    assert rc == abi.ERROR_UNSUPPORTED
    rc, _ = _gpu_decode(src, 4, 4, 4, 1, 0, 18, 0)  # pitch not a multiple of 4
    assert rc == abi.ERROR_INVALID_ARGUMENT
    rc, _ = _gpu_decode(src, 4, 4, 2, 3, 0, 16, 0)  # BC3 writes 4-byte pixels
    assert rc == abi.ERROR_INVALID_ARGUMENT


def test_decoded_texture_draws_like_the_oracle():
    """BC1 mip chain decoded on the GPU straight into the texture the draw
    samples, against the oracle's decode and draw."""
    from cases import bc_texture_levels
    from oracle_runner import oracle_render
    from swiftshader_amd import scenes
    rng = np.random.default_rng(77)
    chain = bc_texture_levels(rng, 128, 1)
    levels = []
    for blocks, dec in chain:
        h, w = dec.shape[:2]
        t = torch.empty((h, w, 4), dtype=torch.uint8, device="cuda:0")
        s = torch.from_numpy(blocks).to("cuda:0")
        desc = abi.BCDesc(s.data_ptr(), t.data_ptr(), w, h, w * 4, 4, 1, 0)
        assert runtime.lib().swr_decode_bc(ctypes.byref(desc), None) == 0
        levels.append(t)
    torch.cuda.synchronize()
    for t, (_, dec) in zip(levels, chain):
        assert np.array_equal(t.cpu().numpy(), dec)
    tex = scenes.TextureData(format=abi.FORMAT_B8G8R8A8_UNORM, levels=[d for _, d in chain], anisotropy_enable=1,
                             max_anisotropy=4.0)
    pos, uv, col = scenes.screen_triangles(rng, 800, 80, 60, area_hi=1200)
    draw = scenes.DrawData(width=80, height=60, positions=pos, colors=col, texcoords=uv,
                           shader=abi.SHADER_MODULATE, texture=tex)
    print("[gpu bc] decoded-texture draw", flush=True)
    dd = runtime.prepare(draw, "cuda:0", texture_tensors=levels)
    runtime.draw(dd)
    torch.cuda.synchronize()
    want = oracle_render(draw)
    assert np.array_equal(dd.targets.color.cpu().numpy(), want[0])
    assert np.array_equal(dd.targets.depth.cpu().numpy().view(np.uint32), want[1].view(np.uint32))
