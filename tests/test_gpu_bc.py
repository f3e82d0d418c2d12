"""swr_decode_bc on the GPU against the oracle, bit for bit: BC1-BC5 in
every alpha / sign mode, BC7 in every mode (and the invalid one) and BC6H in
every mode (and the reserved ones), signed and unsigned; ragged and large
images, pixel strides that leave bytes untouched; and the refused cases
(misaligned or wrongly sized destinations)."""
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


BC6_MODES = [0, 1, 2, 6, 10, 14, 18, 22, 26, 30, 3, 7, 11, 15, 19, 23, 27, 31]  # the last four are reserved


def _bptc_blocks(rng, nb, n, mode):
    """Random BC7 blocks of mode `mode` (8: the invalid all-zero mode byte)
    or BC6H blocks whose mode bits read `mode`."""
    b = rng.integers(0, 256, (nb, 16), dtype=np.uint8)
    if n == 7:
        b[:, 0] = 0 if mode == 8 else (b[:, 0] & np.uint8((0xFF << (mode + 1)) & 0xFF)) | np.uint8(1 << mode)
    elif mode in (0, 1):
        b[:, 0] = (b[:, 0] & np.uint8(0xFC)) | np.uint8(mode)
    else:
        b[:, 0] = (b[:, 0] & np.uint8(0xE0)) | np.uint8(mode)
    return b.ravel()


@pytest.mark.parametrize("mode", range(9))
def test_bc7_matches_the_oracle(mode):
    rng = np.random.default_rng(900 + mode)
    for w, h in SHAPES:
        print(f"[gpu bc] BC7 mode {mode} {w}x{h}", flush=True)
        nb = ((w + 3) // 4) * ((h + 3) // 4)
        src = _bptc_blocks(rng, nb, 7, mode)
        pitch = w * 4 + 8
        want = np.full(pitch * h, 0x5A, np.uint8)
        assert oracle_decode_bc(src, want, w, h, pitch, 4, 7, 0) == 1
        rc, got = _gpu_decode(src, w, h, 4, 7, 0, pitch, 0x5A)
        assert rc == 0 and np.array_equal(got, want), (mode, w, h)


@pytest.mark.parametrize("mode", BC6_MODES)
@pytest.mark.parametrize("signed", [0, 1])
def test_bc6h_matches_the_oracle(mode, signed):
    rng = np.random.default_rng(1000 + mode + 100 * signed)
    flag = 0 if signed else 1
    for w, h in SHAPES:
        print(f"[gpu bc] BC6H mode {mode} signed {signed} {w}x{h}", flush=True)
        nb = ((w + 3) // 4) * ((h + 3) // 4)
        src = _bptc_blocks(rng, nb, 6, mode)
        pitch = w * 8 + 6
        want = np.full(pitch * h, 0x5A, np.uint8)
        assert oracle_decode_bc(src, want, w, h, pitch, 8, 6, flag) == 1
        rc, got = _gpu_decode(src, w, h, 8, 6, flag, pitch, 0x5A)
        assert rc == 0 and np.array_equal(got, want), (mode, signed, w, h)


def test_refused_cases():
    src = np.zeros(64, np.uint8)
    rc, _ = _gpu_decode(src, 4, 4, 4, 6, 0, 32, 0)  # BC6H writes 8-byte pixels
    assert rc == abi.ERROR_INVALID_ARGUMENT
    rc, _ = _gpu_decode(src, 4, 4, 8, 7, 0, 32, 0)  # BC7 writes 4-byte pixels
    assert rc == abi.ERROR_INVALID_ARGUMENT
    rc, _ = _gpu_decode(src, 4, 4, 8, 6, 0, 33, 0)  # odd BC6H pitch
    assert rc == abi.ERROR_INVALID_ARGUMENT
    rc, _ = _gpu_decode(src, 4, 4, 4, 8, 0, 16, 0)  # no BC8
    assert rc == abi.ERROR_INVALID_ARGUMENT
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
